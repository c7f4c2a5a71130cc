// aderdg_capi.hip -- the C-ABI of include/aderdg.h: context, resident device
// state, the step sequence of Euler/main.f90:24-100 and graph replay.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aderdg.h"
#include "euler_kernels.hpp"
#include "operators.hpp"

using namespace aderdg;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(ADERDG_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                         \
    } while (0)

// PNPMTable of Euler/ader_dg.f90:28
const double kPNPM[10] = {1.0, 0.33, 0.17, 0.1, 0.069, 0.045, 0.038, 0.03, 0.02, 0.015};

struct TimedLaunch {
    int part;
    hipEvent_t a, b;
};

}  // namespace

struct aderdg_ctx {
    aderdg_desc desc;
    int N, n, s, V;
    int ncell;
    MeshArgs m;
    Operators ops;
    Operators *ops_dev = nullptr;
    hipStream_t stream = nullptr;
    double *scal = nullptr, *pinned = nullptr;
    double *uh[2] = {nullptr, nullptr}, *lduh = nullptr, *fq = nullptr, *ff = nullptr;
    double *xflux = nullptr, *yflux = nullptr;
    double *lmin_old = nullptr, *lmax_old = nullptr, *lmin = nullptr, *lmax = nullptr, *cell_dt = nullptr,
           *dt_partial = nullptr;
    int *rc = nullptr, *rc_final = nullptr, *list = nullptr, *counters = nullptr;
    int *status[2] = {nullptr, nullptr};
    double *lh[2] = {nullptr, nullptr};
    int cur = 0;
    bool have_uh = false;
    // graph replay, one graph per ping-pong parity
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // timing
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> free_events;
    double ms[ADERDG_TIMING_SLOTS] = {};
    int64_t launches[ADERDG_TIMING_SLOTS] = {};
};

namespace {

StepBufs bufs(aderdg_ctx *c)
{
    const int a = c->cur, z = 1 - c->cur;
    StepBufs b;
    b.op = c->ops_dev;
    b.scal = c->scal;
    b.uh_cur = c->uh[a];
    b.uh_next = c->uh[z];
    b.lduh = c->lduh;
    b.fq = c->fq;
    b.ff = c->ff;
    b.xflux = c->xflux;
    b.yflux = c->yflux;
    b.lmin_old = c->lmin_old;
    b.lmax_old = c->lmax_old;
    b.lmin = c->lmin;
    b.lmax = c->lmax;
    b.cell_dt = c->cell_dt;
    b.dt_partial = c->dt_partial;
    b.rc = c->rc;
    b.rc_final = c->rc_final;
    b.list = c->list;
    b.counters = c->counters;
    b.status_cur = c->status[a];
    b.status_next = c->status[z];
    b.lh_cur = c->lh[a];
    b.lh_next = c->lh[z];
    return b;
}

hipEvent_t get_event(aderdg_ctx *c)
{
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int launch(aderdg_ctx *c, int part, double time = 0.0, double tend = 0.0, int use_host = 0)
{
    const StepBufs b = bufs(c);
    int slot = -1;  // timing slots: dt, predictor, update, spread, subcell, face
    switch (part) {
    case kPartDt: slot = 0; break;
    case kPartPredictor: slot = 1; break;
    case kPartUpdate: slot = 2; break;
    case kPartSpread: slot = 3; break;
    case kPartSubcell: slot = 4; break;
    case kPartFace: slot = 5; break;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing && slot >= 0) {
        e0 = get_event(c);
        e1 = get_event(c);
        if (!e0 || !e1) return fail(ADERDG_EHIP, "hipEventCreate failed");
        HIPCHK(hipEventRecord(e0, c->stream));
    }
    hipError_t e;
    if (part == kPartDt)
        e = launch_dt(c->m, b, c->stream, time, tend, use_host);
    else
        e = launch_part(c->N, c->m, b, c->stream, part);
    if (e != hipSuccess) return fail(ADERDG_EHIP, "kernel launch (part %d): %s", part, hipGetErrorString(e));
    if (e1) {
        HIPCHK(hipEventRecord(e1, c->stream));
        c->pending.push_back({slot, e0, e1});
    }
    return 0;
}

// after a stream synchronisation: fold the event pairs into the totals
void collect_timing(aderdg_ctx *c)
{
    for (auto &t : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            c->ms[t.part] += ms;
            c->launches[t.part] += 1;
        }
        c->free_events.push_back(t.a);
        c->free_events.push_back(t.b);
    }
    c->pending.clear();
}

int sync(aderdg_ctx *c)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    collect_timing(c);
    return 0;
}

// DMR states of Euler/riemann_problem.f90:165-175, PDEPrim2Cons of Euler/pde.f90:29-40
void prim2cons(double g, const double *V, double *Q)
{
    Q[0] = V[0];
    Q[1] = V[0] * V[1];
    Q[2] = V[0] * V[2];
    Q[3] = V[3] / (g - 1.0) + 0.5 * V[0] * (V[1] * V[1] + V[2] * V[2]);
}

template <class T>
int dalloc(T **p, size_t count)
{
    HIPCHK(hipMalloc((void **)p, count * sizeof(T) > 0 ? count * sizeof(T) : 8));
    return 0;
}

int seed(aderdg_ctx *c)
{
    return launch(c, kPartSeed);
}

int step_parts(aderdg_ctx *c)
{
    int r;
    if ((r = launch(c, kPartPredictor))) return r;
    if ((r = launch(c, kPartFace))) return r;
    if ((r = launch(c, kPartUpdate))) return r;
    if ((r = launch(c, kPartSpread))) return r;
    if ((r = launch(c, kPartSubcell))) return r;
    return 0;
}

void destroy_graphs(aderdg_ctx *c)
{
    for (int k = 0; k < 2; k++)
        if (c->graph[k]) {
            (void)hipGraphExecDestroy(c->graph[k]);
            c->graph[k] = nullptr;
        }
}

// One device-resident step (time and tend read from the device), captured per parity.
int build_graphs(aderdg_ctx *c)
{
    const int cur0 = c->cur;
    for (int k = 0; k < 2; k++) {
        c->cur = k;
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int r = launch(c, kPartDt);
        if (!r) r = step_parts(c);
        if (!r) r = launch(c, kPartAdvance);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (r) {
            if (g) (void)hipGraphDestroy(g);
            c->cur = cur0;
            return r;
        }
        if (e != hipSuccess) {
            c->cur = cur0;
            return fail(ADERDG_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        }
        e = hipGraphInstantiate(&c->graph[k], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->cur = cur0;
            return fail(ADERDG_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        }
    }
    c->cur = cur0;
    return 0;
}

}  // namespace

extern "C" {

const char *aderdg_last_error(void) { return g_err.c_str(); }

int aderdg_build_exact(void)
{
#ifdef ADERDG_EXACT
    return 1;
#else
    return 0;
#endif
}

int aderdg_create(const aderdg_desc *d, aderdg_ctx **out)
{
    if (!d || !out) return fail(ADERDG_EINVAL, "aderdg_create: null argument");
    *out = nullptr;
    if (d->pde != ADERDG_PDE_EULER && d->pde != ADERDG_PDE_BN)
        return fail(ADERDG_EINVAL, "aderdg_create: unsupported pde %d", d->pde);
    if (d->pde == ADERDG_PDE_BN && d->bc_mode == ADERDG_BC_DMR)
        return fail(ADERDG_EINVAL, "aderdg_create: the DMR boundaries exist for the Euler system only");
    if (d->pde == ADERDG_PDE_BN && !(d->bn_eos[0] > 1.0 && d->bn_eos[1] > 1.0))
        return fail(ADERDG_EINVAL, "aderdg_create: Baer-Nunziato GAMMA_S and GAMMA_G must exceed 1");
    if (d->N < 1 || d->N > kMaxDegree) return fail(ADERDG_EINVAL, "aderdg_create: N=%d outside 1..%d", d->N, kMaxDegree);
    if (d->imax < 1 || d->jmax < 1 || (long long)d->imax * d->jmax > (1LL << 30))
        return fail(ADERDG_EINVAL, "aderdg_create: bad mesh %d x %d", d->imax, d->jmax);
    if (d->bc_mode < 0 || d->bc_mode > 2) return fail(ADERDG_EINVAL, "aderdg_create: bad bc_mode %d", d->bc_mode);
    if (!(d->xr[0] > d->xl[0]) || !(d->xr[1] > d->xl[1])) return fail(ADERDG_EINVAL, "aderdg_create: empty domain");
    if (d->pde == ADERDG_PDE_EULER && !(d->gamma > 1.0)) return fail(ADERDG_EINVAL, "aderdg_create: gamma must exceed 1");
    if (d->exact != 0) return fail(ADERDG_EINVAL, "aderdg_create: exact must be 0");
    aderdg_ctx *c = new aderdg_ctx();
    c->desc = *d;
    c->N = d->N;
    c->n = d->N + 1;
    c->s = 2 * d->N + 1;
    c->V = d->pde == ADERDG_PDE_BN ? 9 : 4;
    c->ncell = d->imax * d->jmax;
    if (!build_operators(d->N, c->ops)) {
        delete c;
        return fail(ADERDG_ESINGULAR, "aderdg_create: operator construction failed (singular matrix)");
    }
    const int n = c->n, s = c->s;
    const aderdg_ops &o = d->ops;
    if (o.wGPN) std::memcpy(c->ops.w, o.wGPN, n * sizeof(double));
    if (o.Kxi) std::memcpy(c->ops.kxi, o.Kxi, n * n * sizeof(double));
    if (o.iK1) std::memcpy(c->ops.ik1, o.iK1, n * n * sizeof(double));
    if (o.F0) std::memcpy(c->ops.f0, o.F0, n * sizeof(double));
    if (o.FLcoeff) std::memcpy(c->ops.flc, o.FLcoeff, n * sizeof(double));
    if (o.FRcoeff) std::memcpy(c->ops.frc, o.FRcoeff, n * sizeof(double));
    if (o.uh2lim) std::memcpy(c->ops.u2l, o.uh2lim, s * n * sizeof(double));
    if (o.lim2uh) std::memcpy(c->ops.l2u, o.lim2uh, n * s * sizeof(double));
    if (o.uh2lob) std::memcpy(c->ops.u2b, o.uh2lob, n * n * sizeof(double));

    MeshArgs &m = c->m;
    m.pde = d->pde;
    for (int k = 0; k < 4; k++) m.eos[k] = d->bn_eos[k];
    m.imax = d->imax;
    m.jmax = d->jmax;
    m.ncell = c->ncell;
    m.bc_mode = d->bc_mode;
    m.wrap = d->bc_mode == ADERDG_BC_PERIODIC;
    m.dx = (d->xr[0] - d->xl[0]) / (double)d->imax;  // Euler/initialize.f90:43
    m.dy = (d->xr[1] - d->xl[1]) / (double)d->jmax;
    m.gamma = d->gamma;
    m.cflpnpm = d->cfl * kPNPM[d->N];
    const double pi = 4.0 * std::atan(1.0);
    m.dmr_a = 1.0 / 6.0 + 1.0 / std::tan(pi / 3.0);
    m.dmr_b = 10.0 / std::sin(pi / 3.0);
    const double Va[4] = {8.0, 8.25 * std::cos(pi / 6.0), -8.25 * std::sin(pi / 6.0), 116.5};
    const double Vb[4] = {1.4, 0.0, 0.0, 1.0};
    prim2cons(d->gamma, Va, m.qa);
    prim2cons(d->gamma, Vb, m.qb);

    auto bail = [&](int code) {
        aderdg_destroy(c);
        return code;
    };
    if (hipSetDevice(d->device) != hipSuccess) return bail(fail(ADERDG_EHIP, "hipSetDevice(%d) failed", d->device));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(ADERDG_EHIP, "hipStreamCreate failed"));
    const size_t nc = c->ncell, V = c->V, NN = (size_t)n * n, SS = (size_t)s * s;
    int r = 0;
    r = r ? r : dalloc(&c->ops_dev, 1);
    r = r ? r : dalloc(&c->scal, 4);
    r = r ? r : dalloc(&c->uh[0], V * NN * nc);
    r = r ? r : dalloc(&c->uh[1], V * NN * nc);
    r = r ? r : dalloc(&c->lduh, V * NN * nc);
    r = r ? r : dalloc(&c->fq, 4 * V * n * nc);
    r = r ? r : dalloc(&c->ff, 4 * V * n * nc);
    const size_t nside = d->pde == ADERDG_PDE_BN ? 2 : 1;
    r = r ? r : dalloc(&c->xflux, nside * V * n * (size_t)(d->imax + 1) * d->jmax);
    r = r ? r : dalloc(&c->yflux, nside * V * n * (size_t)d->imax * (d->jmax + 1));
    r = r ? r : dalloc(&c->lmin_old, V * nc);
    r = r ? r : dalloc(&c->lmax_old, V * nc);
    r = r ? r : dalloc(&c->lmin, V * nc);
    r = r ? r : dalloc(&c->lmax, V * nc);
    r = r ? r : dalloc(&c->cell_dt, nc);
    r = r ? r : dalloc(&c->dt_partial, (size_t)kDtGrid);
    r = r ? r : dalloc(&c->rc, nc);
    r = r ? r : dalloc(&c->rc_final, nc);
    r = r ? r : dalloc(&c->list, nc);
    r = r ? r : dalloc(&c->counters, (size_t)kNumCounters);
    r = r ? r : dalloc(&c->status[0], nc);
    r = r ? r : dalloc(&c->status[1], nc);
    r = r ? r : dalloc(&c->lh[0], V * SS * nc);
    r = r ? r : dalloc(&c->lh[1], V * SS * nc);
    if (r) return bail(r);
    if (hipHostMalloc((void **)&c->pinned, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return bail(fail(ADERDG_EHIP, "hipHostMalloc failed"));
    if (hipMemcpy(c->ops_dev, &c->ops, sizeof(Operators), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->counters, 0, kNumCounters * sizeof(int)) != hipSuccess ||
        hipMemset(c->scal, 0, 4 * sizeof(double)) != hipSuccess ||
        hipMemset(c->rc, 0, nc * sizeof(int)) != hipSuccess || hipMemset(c->rc_final, 0, nc * sizeof(int)) != hipSuccess ||
        hipMemset(c->status[0], 0, nc * sizeof(int)) != hipSuccess ||
        hipMemset(c->status[1], 0, nc * sizeof(int)) != hipSuccess)
        return bail(fail(ADERDG_EHIP, "device initialisation failed"));
    *out = c;
    return 0;
}

int aderdg_destroy(aderdg_ctx *c)
{
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    destroy_graphs(c);
    collect_timing(c);
    for (auto e : c->free_events) (void)hipEventDestroy(e);
    void *ptrs[] = {c->ops_dev, c->scal, c->uh[0], c->uh[1], c->lduh, c->fq, c->ff, c->xflux, c->yflux, c->lmin_old, c->lmax_old, c->lmin,
                    c->lmax, c->cell_dt, c->dt_partial, c->rc, c->rc_final, c->list, c->counters, c->status[0],
                    c->status[1], c->lh[0], c->lh[1]};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int aderdg_operators(int N, double *xi, double *w, double *kxi, double *dudx, double *ik1, double *f0, double *flc,
                     double *frc, double *u2l, double *l2u, double *u2b)
{
    Operators o;
    if (N < 1 || N > kMaxDegree) return fail(ADERDG_EINVAL, "aderdg_operators: N=%d outside 1..%d", N, kMaxDegree);
    if (!build_operators(N, o)) return fail(ADERDG_ESINGULAR, "aderdg_operators: singular matrix");
    const int n = N + 1, s = 2 * N + 1;
    if (xi) std::memcpy(xi, o.xi, n * sizeof(double));
    if (w) std::memcpy(w, o.w, n * sizeof(double));
    if (kxi) std::memcpy(kxi, o.kxi, n * n * sizeof(double));
    if (dudx) std::memcpy(dudx, o.dudx, n * n * sizeof(double));
    if (ik1) std::memcpy(ik1, o.ik1, n * n * sizeof(double));
    if (f0) std::memcpy(f0, o.f0, n * sizeof(double));
    if (flc) std::memcpy(flc, o.flc, n * sizeof(double));
    if (frc) std::memcpy(frc, o.frc, n * sizeof(double));
    if (u2l) std::memcpy(u2l, o.u2l, s * n * sizeof(double));
    if (l2u) std::memcpy(l2u, o.l2u, n * s * sizeof(double));
    if (u2b) std::memcpy(u2b, o.u2b, n * n * sizeof(double));
    return 0;
}

int aderdg_get_ops(aderdg_ctx *c, double *w, double *kxi, double *ik1, double *f0, double *flc, double *frc,
                   double *u2l, double *l2u, double *u2b)
{
    if (!c) return fail(ADERDG_EINVAL, "aderdg_get_ops: null context");
    const int n = c->n, s = c->s;
    if (w) std::memcpy(w, c->ops.w, n * sizeof(double));
    if (kxi) std::memcpy(kxi, c->ops.kxi, n * n * sizeof(double));
    if (ik1) std::memcpy(ik1, c->ops.ik1, n * n * sizeof(double));
    if (f0) std::memcpy(f0, c->ops.f0, n * sizeof(double));
    if (flc) std::memcpy(flc, c->ops.flc, n * sizeof(double));
    if (frc) std::memcpy(frc, c->ops.frc, n * sizeof(double));
    if (u2l) std::memcpy(u2l, c->ops.u2l, s * n * sizeof(double));
    if (l2u) std::memcpy(l2u, c->ops.l2u, n * s * sizeof(double));
    if (u2b) std::memcpy(u2b, c->ops.u2b, n * n * sizeof(double));
    return 0;
}

int aderdg_upload_uh(aderdg_ctx *c, const double *uh)
{
    if (!c || !uh) return fail(ADERDG_EINVAL, "aderdg_upload_uh: null argument");
    const size_t bytes = sizeof(double) * c->V * c->n * c->n * (size_t)c->ncell;
    HIPCHK(hipMemcpyAsync(c->uh[c->cur], uh, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->status[c->cur], 0, sizeof(int) * c->ncell, c->stream));
    HIPCHK(hipMemsetAsync(c->rc_final, 0, sizeof(int) * c->ncell, c->stream));
    int r = seed(c);
    if (r) return r;
    c->have_uh = true;
    return sync(c);
}

int aderdg_download_uh(aderdg_ctx *c, double *uh)
{
    if (!c || !uh) return fail(ADERDG_EINVAL, "aderdg_download_uh: null argument");
    const size_t bytes = sizeof(double) * c->V * c->n * c->n * (size_t)c->ncell;
    HIPCHK(hipMemcpyAsync(uh, c->uh[c->cur], bytes, hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

int aderdg_set_limiter(aderdg_ctx *c, const int32_t *status, const double *Lh)
{
    if (!c || !status || !Lh) return fail(ADERDG_EINVAL, "aderdg_set_limiter: null argument");
    if (!c->have_uh) return fail(ADERDG_ESTATE, "aderdg_set_limiter: upload uh first");
    const size_t nc = c->ncell, blk = (size_t)c->V * c->s * c->s;
    std::vector<int32_t> st(nc);
    for (size_t k = 0; k < nc; k++) st[k] = status[k] ? 1 : 0;
    HIPCHK(hipMemcpyAsync(c->status[c->cur], st.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->lh[c->cur], Lh, nc * blk * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int r = seed(c);
    if (r) return r;
    return sync(c);
}

int aderdg_get_limiter(aderdg_ctx *c, int32_t *status, int32_t *recompute, double *Lh, double *lmin, double *lmax)
{
    if (!c) return fail(ADERDG_EINVAL, "aderdg_get_limiter: null context");
    const size_t nc = c->ncell, blk = (size_t)c->V * c->s * c->s, V = c->V;
    std::vector<int32_t> st(nc);
    HIPCHK(hipMemcpyAsync(st.data(), c->status[c->cur], nc * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    if (recompute) HIPCHK(hipMemcpyAsync(recompute, c->rc_final, nc * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    if (Lh) HIPCHK(hipMemcpyAsync(Lh, c->lh[c->cur], nc * blk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    std::vector<double> mn, mx;
    if (lmin) {
        mn.resize(V * nc);
        HIPCHK(hipMemcpyAsync(mn.data(), c->lmin, V * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    if (lmax) {
        mx.resize(V * nc);
        HIPCHK(hipMemcpyAsync(mx.data(), c->lmax, V * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    int r = sync(c);
    if (r) return r;
    for (size_t k = 0; k < nc; k++) {
        if (status) status[k] = st[k];
        if (Lh && !st[k]) std::memset(Lh + k * blk, 0, blk * sizeof(double));
        for (size_t v = 0; v < V; v++) {
            if (lmin) lmin[v + V * k] = mn[v * nc + k];
            if (lmax) lmax[v + V * k] = mx[v * nc + k];
        }
    }
    return 0;
}

int aderdg_calc_time_step(aderdg_ctx *c, double time, double tend, double *dt)
{
    if (!c) return fail(ADERDG_EINVAL, "aderdg_calc_time_step: null context");
    if (!c->have_uh) return fail(ADERDG_ESTATE, "aderdg_calc_time_step: no solution uploaded");
    int r = launch(c, kPartDt, time, tend, 1);
    if (r) return r;
    if (dt) {
        HIPCHK(hipMemcpyAsync(c->pinned, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        if ((r = sync(c))) return r;
        *dt = c->pinned[0];
    }
    return 0;
}

int aderdg_limiter_begin(aderdg_ctx *c)
{
    if (!c) return fail(ADERDG_EINVAL, "aderdg_limiter_begin: null context");
    if (!c->have_uh) return fail(ADERDG_ESTATE, "aderdg_limiter_begin: no solution uploaded");
    return 0;
}

int aderdg_space_time_predictor(aderdg_ctx *c)
{
    if (!c) return fail(ADERDG_EINVAL, "aderdg_space_time_predictor: null context");
    if (!c->have_uh) return fail(ADERDG_ESTATE, "aderdg_space_time_predictor: no solution uploaded");
    return launch(c, kPartPredictor);
}

int aderdg_riemann_surface_update(aderdg_ctx *c, int32_t *nrec)
{
    if (!c) return fail(ADERDG_EINVAL, "aderdg_riemann_surface_update: null context");
    int r = launch(c, kPartFace);
    if (!r) r = launch(c, kPartUpdate);
    if (r) return r;
    if (nrec) {
        HIPCHK(hipMemcpyAsync(c->pinned + 4, c->counters, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        if ((r = sync(c))) return r;
        int tmp[2];
        std::memcpy(tmp, c->pinned + 4, sizeof tmp);
        *nrec = tmp[kCntTroubled];
    }
    return 0;
}

int aderdg_limiter_end(aderdg_ctx *c)
{
    if (!c) return fail(ADERDG_EINVAL, "aderdg_limiter_end: null context");
    int r = launch(c, kPartSpread);
    if (!r) r = launch(c, kPartSubcell);
    if (r) return r;
    c->cur = 1 - c->cur;  // update_limiter: Lh = NewLh, oldstatus = status, uh <- new
    return 0;
}

int aderdg_step(aderdg_ctx *c, double time, double tend, double *dt, int32_t *nrec)
{
    int r = aderdg_calc_time_step(c, time, tend, dt);
    if (!r) r = aderdg_limiter_begin(c);
    if (!r) r = aderdg_space_time_predictor(c);
    if (!r) r = aderdg_riemann_surface_update(c, nrec);
    if (!r) r = aderdg_limiter_end(c);
    return r;
}

int aderdg_run(aderdg_ctx *c, int nsteps, double time, double tend, double *time_out, double *dt_out,
               int32_t *nrec_out)
{
    if (!c) return fail(ADERDG_EINVAL, "aderdg_run: null context");
    if (nsteps < 0) return fail(ADERDG_EINVAL, "aderdg_run: negative step count");
    if (!c->have_uh) return fail(ADERDG_ESTATE, "aderdg_run: no solution uploaded");
    c->pinned[0] = 0.0;
    c->pinned[1] = time;
    c->pinned[2] = tend;
    HIPCHK(hipMemcpyAsync(c->scal, c->pinned, 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int r = sync(c);
    if (r) return r;
    if (!c->timing && !c->graph[0]) {
        if ((r = build_graphs(c))) return r;
    }
    for (int k = 0; k < nsteps; k++) {
        if (c->timing) {
            r = launch(c, kPartDt);
            if (!r) r = step_parts(c);
            if (!r) r = launch(c, kPartAdvance);
            if (r) return r;
        } else {
            HIPCHK(hipGraphLaunch(c->graph[c->cur], c->stream));
        }
        c->cur = 1 - c->cur;
    }
    HIPCHK(hipMemcpyAsync(c->pinned, c->scal, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->pinned + 4, c->counters, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if ((r = sync(c))) return r;
    if (time_out) *time_out = c->pinned[1];
    if (dt_out) *dt_out = c->pinned[0];
    if (nrec_out) {
        int tmp[2];
        std::memcpy(tmp, c->pinned + 4, sizeof tmp);
        *nrec_out = tmp[kCntTroubled];
    }
    return 0;
}

namespace {
int predictor_cells(aderdg_ctx *c, int ncells, double dt, const double *luh, double *lqhi, double *lFhi, double *lShi,
                    double *lqbnd, double *lFbnd)
{
    const size_t n = c->n, V = c->V, nc = ncells;
    double *d_uh = nullptr, *d_q = nullptr, *d_f = nullptr, *d_qb = nullptr, *d_fb = nullptr, *d_s = nullptr;
    int r = dalloc(&d_uh, V * n * n * nc);
    r = r ? r : dalloc(&d_q, V * n * n * nc);
    r = r ? r : dalloc(&d_f, 2 * V * n * n * nc);
    r = r ? r : dalloc(&d_qb, 4 * V * n * nc);
    r = r ? r : dalloc(&d_fb, 4 * V * n * nc);
    r = r ? r : dalloc(&d_s, V * n * n * nc);
    if (!r) {
        hipError_t e = hipMemcpyAsync(d_uh, luh, V * n * n * nc * sizeof(double), hipMemcpyHostToDevice, c->stream);
        c->pinned[0] = dt;
        if (e == hipSuccess) e = hipMemcpyAsync(c->scal, c->pinned, sizeof(double), hipMemcpyHostToDevice, c->stream);
        MeshArgs m = c->m;
        m.ncell = ncells;
        TestBufs t{d_uh, d_q, d_f, d_qb, d_fb, d_s};
        if (e == hipSuccess) e = launch_test_predictor(c->N, m, bufs(c), t, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(lqhi, d_q, V * n * n * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(lFhi, d_f, 2 * V * n * n * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(lqbnd, d_qb, 4 * V * n * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(lFbnd, d_fb, 4 * V * n * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && lShi)
            e = hipMemcpyAsync(lShi, d_s, V * n * n * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) r = fail(ADERDG_EHIP, "aderdg_predictor_cells: %s", hipGetErrorString(e));
    }
    void *ptrs[] = {d_uh, d_q, d_f, d_qb, d_fb, d_s};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    return r;
}
}  // namespace

int aderdg_predictor_cells(aderdg_ctx *c, int ncells, double dt, const double *luh, double *lqhi, double *lFhi,
                           double *lqbnd, double *lFbnd)
{
    if (!c || !luh || !lqhi || !lFhi || !lqbnd || !lFbnd || ncells < 1)
        return fail(ADERDG_EINVAL, "aderdg_predictor_cells: bad argument");
    if (c->desc.pde != ADERDG_PDE_EULER) return fail(ADERDG_EINVAL, "aderdg_predictor_cells: use the _bn form");
    return predictor_cells(c, ncells, dt, luh, lqhi, lFhi, nullptr, lqbnd, lFbnd);
}

int aderdg_predictor_cells_bn(aderdg_ctx *c, int ncells, double dt, const double *luh, double *lqhi, double *lFhi,
                              double *lShi, double *lqbnd, double *lFbnd)
{
    if (!c || !luh || !lqhi || !lFhi || !lShi || !lqbnd || !lFbnd || ncells < 1)
        return fail(ADERDG_EINVAL, "aderdg_predictor_cells_bn: bad argument");
    if (c->desc.pde != ADERDG_PDE_BN) return fail(ADERDG_EINVAL, "aderdg_predictor_cells_bn: not a Baer-Nunziato context");
    return predictor_cells(c, ncells, dt, luh, lqhi, lFhi, lShi, lqbnd, lFbnd);
}

int aderdg_timing_enable(aderdg_ctx *c, int on)
{
    if (!c) return fail(ADERDG_EINVAL, "aderdg_timing_enable: null context");
    int r = sync(c);
    if (r) return r;
    c->timing = on != 0;
    for (int k = 0; k < ADERDG_TIMING_SLOTS; k++) {
        c->ms[k] = 0.0;
        c->launches[k] = 0;
    }
    return 0;
}

int aderdg_timing_get(aderdg_ctx *c, double ms[ADERDG_TIMING_SLOTS], int64_t launches[ADERDG_TIMING_SLOTS])
{
    if (!c) return fail(ADERDG_EINVAL, "aderdg_timing_get: null context");
    int r = sync(c);
    if (r) return r;
    for (int k = 0; k < ADERDG_TIMING_SLOTS; k++) {
        if (ms) ms[k] = c->ms[k];
        if (launches) launches[k] = c->launches[k];
    }
    return 0;
}

// Operation counts per cell-step with the oracle's convention (+ - * / sqrt = 1).
// Predictor launch: Euler/space_time_predictor.f90 + Euler/volume_integral.f90
// (Baer-Nunziato: + gradients, NCP source and its time average).
double aderdg_flops_predictor(int pde, int N)
{
    if ((pde != ADERDG_PDE_EULER && pde != ADERDG_PDE_BN) || N < 1) return 0.0;
    const bool bn = pde == ADERDG_PDE_BN;
    const double n = N + 1, V = bn ? 9 : 4, cF = bn ? 76 : 26, cNcp = 34;
    const double init = 3 * V * n * n * n;
    double per_slice = n * n * cF + 2 * (3 * n) + 2 * V * n * n * (2 * n + 2);
    if (bn) per_slice += 2 * V * n * n * (2 * n + 1) + n * n * (3 + 2 * V) + n * n * cNcp;
    const double time_contraction = n * n * (2 + V * n * (2 * n + 1));
    const double picard = n * (n * per_slice + time_contraction);
    const double averages = (bn ? 4 : 3) * V * n * n * 2 * n;
    const double extrapolation = 8 * V * n * 2 * n;
    const double volume = 2 * V * n * n * (2 * n + 3) + (bn ? 3 * V * n * n : 0);
    return init + picard + averages + extrapolation + volume;
}

// Update launch: Riemann (one x- and one y-face line per cell), surface integral,
// element update, DMP and the next step's time-step candidates.
double aderdg_flops_update(int pde, int N)
{
    if ((pde != ADERDG_PDE_EULER && pde != ADERDG_PDE_BN) || N < 1) return 0.0;
    const bool bn = pde == ADERDG_PDE_BN;
    const double n = N + 1, V = bn ? 9 : 4, s = 2 * N + 1;
    const double eig = bn ? 48 : 17;
    const double rusanov = bn ? 2 * eig + 3 * (27 + 14 + 27) + 9 * 6 + 9 * 3 + 9 * 4 : 2 * eig + 4 * 6;
    const double faces = 2 * n * rusanov;
    const double surface = 2 * V * n * (1 + 5 * n);
    const double update = 4 * V * n * n;
    const double positivity = bn ? 32 : 9;
    const double dmp = 2 * V * n * s * (n + s) + 4 * V * n * n * n + V * 4 + s * s * positivity;
    const double dt = n * n * (2 * eig + 5);
    return faces + surface + update + dmp + dt;
}

}  // extern "C"
