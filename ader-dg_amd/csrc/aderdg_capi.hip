// aderdg_capi.hip -- the C-ABI of include/aderdg.h for one solve: context,
// resident device state, the step sequence of Euler/main.f90:24-100 and graph
// replay.  The context is a solve group (group.hpp) holding a single solve.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/aderdg.h"
#include "capi_common.hpp"
#include "group.hpp"
#include "operators.hpp"
#include "percell.hpp"
#include "step_kernels.hpp"

using namespace aderdg;

namespace {

thread_local std::string g_err;

// PNPMTable of Euler/ader_dg.f90:28
const double kPNPM[10] = {1.0, 0.33, 0.17, 0.1, 0.069, 0.045, 0.038, 0.03, 0.02, 0.015};

struct TimedLaunch {
    int part;
    hipEvent_t a, b;
};

// Staging of aderdg_predictor_cells_noalloc: a file-scope device array of 16 MiB, one copy on
// every GPU the library runs on, held until the process ends.  In host code its name is only a
// handle: the array's address on a device comes from hipGetSymbolAddress, resolved once per
// device by scratch_resolve (called from aderdg_create with the context's device current) and
// kept in g_scratch.  Every no-allocation call on a GPU shares that GPU's copy.
constexpr size_t kScratchDoubles = (size_t)2 << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxScratchDevices] = {};

int scratch_resolve(int device);

// DMR states of Euler/riemann_problem.f90:165-175, PDEPrim2Cons of Euler/pde.f90:29-40
void prim2cons(double g, const double *V, double *Q)
{
    Q[0] = V[0];
    Q[1] = V[0] * V[1];
    Q[2] = V[0] * V[2];
    Q[3] = V[3] / (g - 1.0) + 0.5 * V[0] * (V[1] * V[1] + V[2] * V[2]);
}

}  // namespace

namespace aderdg {

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

}  // namespace aderdg

namespace {
int scratch_resolve(int device)
{
    if (device < 0 || device >= kMaxScratchDevices)
        return aderdg::fail(ADERDG_EINVAL, "device %d outside 0..%d", device, kMaxScratchDevices - 1);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return aderdg::fail(ADERDG_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch[device] = static_cast<double *>(p);
    return 0;
}
}  // namespace

namespace aderdg {

int check_desc(const aderdg_desc *d, const char *who)
{
    if (!d) return fail(ADERDG_EINVAL, "%s: null descriptor", who);
    if (d->pde != ADERDG_PDE_EULER && d->pde != ADERDG_PDE_BN && d->pde != ADERDG_PDE_AD)
        return fail(ADERDG_EINVAL, "%s: unsupported pde %d", who, d->pde);
    if (d->pde == ADERDG_PDE_AD) {
        for (int k = 0; k < 4; k++)
            if (d->ad_bc[k] < 1 || d->ad_bc[k] > 3)
                return fail(ADERDG_EINVAL, "%s: advection-diffusion boundary code %d outside 1..3", who, d->ad_bc[k]);
        if (!(d->ad_a >= 0.0)) return fail(ADERDG_EINVAL, "%s: negative diffusivity", who);
        if (d->ad_fixed_predictor != 0 && d->ad_fixed_predictor != 1)
            return fail(ADERDG_EINVAL, "%s: ad_fixed_predictor must be 0 or 1", who);
    }
    if (d->pde == ADERDG_PDE_BN && d->bc_mode == ADERDG_BC_DMR)
        return fail(ADERDG_EINVAL, "%s: the DMR boundaries exist for the Euler system only", who);
    if (d->pde == ADERDG_PDE_BN && !(d->bn_eos[0] > 1.0 && d->bn_eos[1] > 1.0))
        return fail(ADERDG_EINVAL, "%s: Baer-Nunziato GAMMA_S and GAMMA_G must exceed 1", who);
    if (d->N < 1 || d->N > kMaxDegree) return fail(ADERDG_EINVAL, "%s: N=%d outside 1..%d", who, d->N, kMaxDegree);
    if (d->imax < 1 || d->jmax < 1 || (long long)d->imax * d->jmax > (1LL << 28))
        return fail(ADERDG_EINVAL, "%s: bad mesh %d x %d", who, d->imax, d->jmax);
    if (d->bc_mode < 0 || d->bc_mode > 3) return fail(ADERDG_EINVAL, "%s: bad bc_mode %d", who, d->bc_mode);
    if (d->pde != ADERDG_PDE_AD && d->bc_mode == ADERDG_BC_SIDES) {
        for (int k = 0; k < 4; k++)
            if (d->bc[k] < ADERDG_BC_INFLOW || d->bc[k] > ADERDG_BC_WRAP)
                return fail(ADERDG_EINVAL, "%s: boundary code %d of side %d outside 0..3", who, d->bc[k], k);
        if ((d->bc[0] == ADERDG_BC_WRAP) != (d->bc[1] == ADERDG_BC_WRAP) ||
            (d->bc[2] == ADERDG_BC_WRAP) != (d->bc[3] == ADERDG_BC_WRAP))
            return fail(ADERDG_EINVAL, "%s: periodic boundaries come in pairs (codes %d %d %d %d)", who, d->bc[0],
                        d->bc[1], d->bc[2], d->bc[3]);
        const int V = d->pde == ADERDG_PDE_BN ? 9 : 4;
        for (int k = 0; k < 4; k++)
            if (d->bc[k] == ADERDG_BC_INFLOW)
                for (int v = 0; v < V; v++)
                    if (!std::isfinite(d->bc_inflow[v]))
                        return fail(ADERDG_EINVAL, "%s: inflow state component %d is not finite", who, v);
    }
    if (!(d->xr[0] > d->xl[0]) || !(d->xr[1] > d->xl[1])) return fail(ADERDG_EINVAL, "%s: empty domain", who);
    if (d->pde == ADERDG_PDE_EULER && !(d->gamma > 1.0)) return fail(ADERDG_EINVAL, "%s: gamma must exceed 1", who);
    if (!(d->cfl > 0.0)) return fail(ADERDG_EINVAL, "%s: cfl must be positive", who);
    if (d->exact != 0) return fail(ADERDG_EINVAL, "%s: exact must be 0", who);
    return 0;
}

MeshArgs mesh_from_desc(const aderdg_desc *d)
{
    MeshArgs m{};
    m.pde = d->pde;
    for (int k = 0; k < 4; k++) m.eos[k] = d->bn_eos[k];
    m.imax = d->imax;
    m.jmax = d->jmax;
    m.ncell = d->imax * d->jmax;
    m.bc_mode = d->bc_mode;
    for (int k = 0; k < 4; k++)
        m.side[k] = d->bc_mode == ADERDG_BC_SIDES      ? d->bc[k]
                    : d->bc_mode == ADERDG_BC_PERIODIC ? kSideWrap
                                                       : kSideTransmit;
    m.wrap = (m.side[0] == kSideWrap ? 1 : 0) | (m.side[2] == kSideWrap ? 2 : 0);
    if (d->bc_mode == ADERDG_BC_DMR) m.wrap = 0;
    for (int v = 0; v < 9; v++) m.qin[v] = d->bc_inflow[v];
    m.xl[0] = d->xl[0];
    m.xl[1] = d->xl[1];
    m.adv[0] = d->ad_c[0];
    m.adv[1] = d->ad_c[1];
    m.diff = d->ad_a;
    m.ad_fixed = d->ad_fixed_predictor;
    for (int k = 0; k < 4; k++) m.ad_bc[k] = d->ad_bc[k];
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
    return m;
}

int ops_from_desc(const aderdg_desc *d, Operators *o, const char *who)
{
    if (!build_operators(d->N, *o)) return fail(ADERDG_ESINGULAR, "%s: operator construction failed (singular matrix)", who);
    const int n = d->N + 1, s = 2 * d->N + 1;
    const aderdg_ops &h = d->ops;
    if (h.wGPN) std::memcpy(o->w, h.wGPN, n * sizeof(double));
    if (h.Kxi) std::memcpy(o->kxi, h.Kxi, n * n * sizeof(double));
    if (h.iK1) std::memcpy(o->ik1, h.iK1, n * n * sizeof(double));
    if (h.F0) std::memcpy(o->f0, h.F0, n * sizeof(double));
    if (h.FLcoeff) std::memcpy(o->flc, h.FLcoeff, n * sizeof(double));
    if (h.FRcoeff) std::memcpy(o->frc, h.FRcoeff, n * sizeof(double));
    if (h.uh2lim) std::memcpy(o->u2l, h.uh2lim, s * n * sizeof(double));
    if (h.lim2uh) std::memcpy(o->l2u, h.lim2uh, n * s * sizeof(double));
    if (h.uh2lob) std::memcpy(o->u2b, h.uh2lob, n * n * sizeof(double));
    return 0;
}

}  // namespace aderdg

struct aderdg_ctx {
    aderdg_desc desc;
    int N, n, s, V;
    int ncell;
    hipStream_t stream = nullptr;
    Group g;
    int slot = 0;
    SolveState st{};  // host view of the solve's state; parity mirrors the device
    double *pinned = nullptr;
    bool have_uh = false;
    hipGraphExec_t graph = nullptr;  // one device-resident step
    // timing
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> free_events;
    double ms[ADERDG_TIMING_SLOTS] = {};
    int64_t launches[ADERDG_TIMING_SLOTS] = {};
};

namespace {

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

int timing_slot(int part)
{
    switch (part) {  // dt, predictor, update, spread, subcell, face, ranges
    case kPartDt: return 0;
    case kPartPredictor: return 1;
    case kPartUpdate: return 2;
    case kPartSpread: return 3;
    case kPartSubcell: return 4;
    case kPartFace: return 5;
    case kPartRanges: return 6;
    default: return -1;
    }
}

int launch(aderdg_ctx *c, int part, double time = 0.0, double tend = 0.0, int use_host = 0)
{
    const int slot = timing_slot(part);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing && slot >= 0) {
        e0 = get_event(c);
        e1 = get_event(c);
        if (!e0 || !e1) return fail(ADERDG_EHIP, "hipEventCreate failed");
        HIPCHK(hipEventRecord(e0, c->stream));
    }
    hipError_t e;
    if (part == kPartDt)
        e = launch_dt(c->g.b, c->stream, time, tend, use_host, c->g.dt_grid());
    else
        e = c->g.launch(part);
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

// Per-cell limiter seeds and time-step candidates of the current solution.
int seed(aderdg_ctx *c)
{
    c->st.seed = 1;
    c->st.active = 1;
    c->st.dtmin = ~0ull;
    c->st.max_steps = INT_MAX;
    c->st.stop_at_tend = 0;
    int r = c->g.write_state(c->slot, c->st);
    if (r) return fail(r, "aderdg: state upload failed");
    if ((r = launch(c, kPartSeed))) return r;
    if ((r = launch(c, kPartSeedDone))) return r;
    c->st.seed = 0;
    return 0;
}

int step_parts(aderdg_ctx *c)
{
    int r;
    if (c->g.pde != kPdeAd && (r = launch(c, kPartRanges))) return r;  // no limiter, no ranges
    if ((r = launch(c, kPartPredictor))) return r;
    if ((r = launch(c, kPartFace))) return r;
    if ((r = launch(c, kPartUpdate))) return r;
    if ((r = launch(c, kPartSpread))) return r;
    if ((r = launch(c, kPartSubcell))) return r;
    return 0;
}

// One device-resident step (time, tend, parity read from the device).
int build_graph(aderdg_ctx *c)
{
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = launch(c, kPartDt);
    if (!r) r = step_parts(c);
    if (!r) r = launch(c, kPartAdvance);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return fail(ADERDG_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(ADERDG_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
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
    int r = check_desc(d, "aderdg_create");
    if (r) return r;
    aderdg_ctx *c = new aderdg_ctx();
    c->desc = *d;
    c->N = d->N;
    c->n = d->N + 1;
    c->s = 2 * d->N + 1;
    c->V = d->pde == ADERDG_PDE_BN ? 9 : (d->pde == ADERDG_PDE_AD ? 1 : 4);
    c->ncell = d->imax * d->jmax;
    Operators ops;
    if ((r = ops_from_desc(d, &ops, "aderdg_create"))) {
        delete c;
        return r;
    }
    const MeshArgs m = mesh_from_desc(d);
    auto bail = [&](int code) {
        aderdg_destroy(c);
        return code;
    };
    if (hipSetDevice(d->device) != hipSuccess) return bail(fail(ADERDG_EHIP, "hipSetDevice(%d) failed", d->device));
    if ((r = scratch_resolve(d->device))) return bail(r);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(ADERDG_EHIP, "hipStreamCreate failed"));
    const GroupCaps caps{c->ncell, (int64_t)(d->imax + 1) * d->jmax, (int64_t)d->imax * (d->jmax + 1), 1};
    if ((r = c->g.init(d->pde, d->N, caps, ops, 0, c->stream))) return bail(fail(r, "aderdg_create: device allocation failed"));
    SolveState st0{};
    st0.dtmin = ~0ull;
    st0.max_steps = INT_MAX;
    if ((r = c->g.add(m, st0, &c->slot))) return bail(fail(r, "aderdg_create: solve placement failed"));
    c->st = st0;
    if (hipHostMalloc((void **)&c->pinned, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return bail(fail(ADERDG_EHIP, "hipHostMalloc failed"));
    *out = c;
    return 0;
}

int aderdg_destroy(aderdg_ctx *c)
{
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    collect_timing(c);
    for (auto e : c->free_events) (void)hipEventDestroy(e);
    c->g.release();
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
    const Operators &o = c->g.ops;
    if (w) std::memcpy(w, o.w, n * sizeof(double));
    if (kxi) std::memcpy(kxi, o.kxi, n * n * sizeof(double));
    if (ik1) std::memcpy(ik1, o.ik1, n * n * sizeof(double));
    if (f0) std::memcpy(f0, o.f0, n * sizeof(double));
    if (flc) std::memcpy(flc, o.flc, n * sizeof(double));
    if (frc) std::memcpy(frc, o.frc, n * sizeof(double));
    if (u2l) std::memcpy(u2l, o.u2l, s * n * sizeof(double));
    if (l2u) std::memcpy(l2u, o.l2u, n * s * sizeof(double));
    if (u2b) std::memcpy(u2b, o.u2b, n * n * sizeof(double));
    return 0;
}

int aderdg_upload_uh(aderdg_ctx *c, const double *uh)
{
    if (!c || !uh) return fail(ADERDG_EINVAL, "aderdg_upload_uh: null argument");
    const int par = c->st.parity;
    int r = c->g.upload_uh(c->slot, uh, par);
    if (r) return fail(r, "aderdg_upload_uh: copy failed");
    HIPCHK(hipMemsetAsync(c->g.b.status[par], 0, sizeof(int) * c->ncell, c->stream));
    HIPCHK(hipMemsetAsync(c->g.b.rc_final, 0, sizeof(int) * c->ncell, c->stream));
    if ((r = seed(c))) return r;
    c->have_uh = true;
    return sync(c);
}

int aderdg_download_uh(aderdg_ctx *c, double *uh)
{
    if (!c || !uh) return fail(ADERDG_EINVAL, "aderdg_download_uh: null argument");
    int r = c->g.download_uh(c->slot, uh, c->st.parity);
    if (r) return fail(r, "aderdg_download_uh: copy failed");
    return sync(c);
}

int aderdg_set_limiter(aderdg_ctx *c, const int32_t *status, const double *Lh)
{
    if (!c || !status || !Lh) return fail(ADERDG_EINVAL, "aderdg_set_limiter: null argument");
    if (!c->have_uh) return fail(ADERDG_ESTATE, "aderdg_set_limiter: upload uh first");
    const size_t nc = c->ncell, blk = (size_t)c->V * c->s * c->s;
    const int par = c->st.parity;
    // limited cells take the pool slots of the solve in cell order
    const SolveArgs &sa = c->g.args(c->slot);
    std::vector<int32_t> st(nc), slot(nc, -1);
    std::vector<double> pool;
    for (size_t k = 0; k < nc; k++) {
        st[k] = status[k] ? 1 : 0;
        if (!st[k]) continue;
        slot[k] = sa.lh0 + (int)(pool.size() / blk);
        pool.insert(pool.end(), Lh + k * blk, Lh + (k + 1) * blk);
    }
    if (pool.size() > (size_t)sa.lhcap * blk)
        return fail(ADERDG_ECAPACITY, "aderdg_set_limiter: %zu limited cells, pool of %d", pool.size() / blk, sa.lhcap);
    HIPCHK(hipMemcpyAsync(c->g.b.status[par], st.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->g.b.lh_slot[par], slot.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    if (!pool.empty())
        HIPCHK(hipMemcpyAsync(c->g.b.lh[par] + (size_t)sa.lh0 * blk, pool.data(), pool.size() * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int r = seed(c);
    if (r) return r;
    return sync(c);
}

int aderdg_seed_limiter(aderdg_ctx *c, const aderdg_ic *ic)
{
    if (!c || !ic) return fail(ADERDG_EINVAL, "aderdg_seed_limiter: null argument");
    if (!c->have_uh) return fail(ADERDG_ESTATE, "aderdg_seed_limiter: upload uh first");
    if (c->desc.pde == ADERDG_PDE_AD) return fail(ADERDG_EINVAL, "aderdg_seed_limiter: advection-diffusion has no limiter");
    if (!ic_fits(c->desc.pde, ic->kind)) return fail(ADERDG_EINVAL, "aderdg_seed_limiter: formula %d does not fit the system", ic->kind);
    IcSpec s{};
    s.kind = ic->kind;
    for (int k = 0; k < kIcParams; k++) s.p[k] = ic->p[k];
    HIPCHK(launch_seed_ic(c->desc.pde, c->N, c->g.b, s, c->slot, (int)c->ncell, c->stream));
    int r = seed(c);
    if (r) return r;
    return sync(c);
}

int aderdg_get_limiter(aderdg_ctx *c, int32_t *status, int32_t *recompute, double *Lh, double *lmin, double *lmax)
{
    if (!c) return fail(ADERDG_EINVAL, "aderdg_get_limiter: null context");
    const size_t nc = c->ncell, blk = (size_t)c->V * c->s * c->s, V = c->V;
    const int par = c->st.parity;
    const SolveArgs &sa = c->g.args(c->slot);
    std::vector<int32_t> st(nc), slot;
    std::vector<double> pool;
    HIPCHK(hipMemcpyAsync(st.data(), c->g.b.status[par], nc * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    if (recompute)
        HIPCHK(hipMemcpyAsync(recompute, c->g.b.rc_final, nc * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    if (Lh && sa.lhcap > 0) {  // the solve's pool segment and the slots of its limited cells
        slot.resize(nc);
        pool.resize((size_t)sa.lhcap * blk);
        HIPCHK(hipMemcpyAsync(slot.data(), c->g.b.lh_slot[par], nc * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(pool.data(), c->g.b.lh[par] + (size_t)sa.lh0 * blk, pool.size() * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
    }
    std::vector<double> mn, mx;
    if (lmin) {
        mn.resize(V * nc);
        HIPCHK(hipMemcpyAsync(mn.data(), c->g.b.lmin, V * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    if (lmax) {
        mx.resize(V * nc);
        HIPCHK(hipMemcpyAsync(mx.data(), c->g.b.lmax, V * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    int r = sync(c);
    if (r) return r;
    for (size_t k = 0; k < nc; k++) {
        if (status) status[k] = st[k];
        if (Lh) {
            const int ls = st[k] && !slot.empty() ? slot[k] - sa.lh0 : -1;
            if (ls >= 0 && ls < sa.lhcap)
                std::memcpy(Lh + k * blk, pool.data() + (size_t)ls * blk, blk * sizeof(double));
            else
                std::memset(Lh + k * blk, 0, blk * sizeof(double));
        }
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
        HIPCHK(hipMemcpyAsync(c->pinned, &c->g.b.state[c->slot].dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
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
    const int r = c->g.pde == kPdeAd ? 0 : launch(c, kPartRanges);
    return r ? r : launch(c, kPartPredictor);
}

int aderdg_riemann_surface_update(aderdg_ctx *c, int32_t *nrec)
{
    if (!c) return fail(ADERDG_EINVAL, "aderdg_riemann_surface_update: null context");
    int r = launch(c, kPartFace);
    if (!r) r = launch(c, kPartUpdate);
    if (r) return r;
    if (nrec) {
        HIPCHK(hipMemcpyAsync(c->pinned + 4, c->g.b.counters, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
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
    if (!r) r = launch(c, kPartFlip);  // update_limiter: Lh = NewLh, oldstatus = status, uh <- new
    if (r) return r;
    c->st.parity ^= 1;
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
    c->st.dt = 0.0;
    c->st.time = time;
    c->st.tend = tend;
    c->st.dtmin = ~0ull;
    c->st.active = 1;
    c->st.steps = 0;
    int r = c->g.write_state(c->slot, c->st);
    if (r) return fail(r, "aderdg_run: state upload failed");
    if ((r = sync(c))) return r;
    if (!c->timing && !c->graph) {
        if ((r = build_graph(c))) return r;
    }
    for (int k = 0; k < nsteps; k++) {
        if (c->timing) {
            r = launch(c, kPartDt);
            if (!r) r = step_parts(c);
            if (!r) r = launch(c, kPartAdvance);
            if (r) return r;
        } else {
            HIPCHK(hipGraphLaunch(c->graph, c->stream));
        }
        c->st.parity ^= 1;
    }
    SolveState *hs = reinterpret_cast<SolveState *>(c->pinned);
    static_assert(sizeof(SolveState) <= 4 * sizeof(double) + 8 * sizeof(int), "pinned staging too small");
    HIPCHK(hipMemcpyAsync(hs, &c->g.b.state[c->slot], sizeof(SolveState), hipMemcpyDeviceToHost, c->stream));
    int cnt[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(cnt, c->g.b.counters, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if ((r = sync(c))) return r;
    c->st.time = hs->time;
    c->st.dt = hs->dt;
    c->st.tend = hs->tend;
    if (time_out) *time_out = hs->time;
    if (dt_out) *dt_out = hs->dt;
    if (nrec_out) *nrec_out = cnt[kCntTroubled];
    return 0;
}

namespace {
int dalloc(double **p, size_t count)
{
    HIPCHK(hipMalloc((void **)p, count > 0 ? count * sizeof(double) : 8));
    return 0;
}

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
        MeshArgs m = c->g.mesh(c->slot);
        m.ncell = ncells;
        TestBufs t{d_uh, d_q, d_f, d_qb, d_fb, d_s};
        if (e == hipSuccess) e = launch_test_predictor(c->N, c->g.variant, m, c->g.b, t, dt, c->stream);
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

int aderdg_predictor_cells_noalloc(aderdg_ctx *c, int ncells, double dt, const double *luh, double *lqhi, double *lFhi,
                                   double *lShi, double *lqbnd, double *lFbnd)
{
    if (!c || !luh || !lqhi || !lFhi || !lqbnd || !lFbnd || ncells < 1)
        return fail(ADERDG_EINVAL, "aderdg_predictor_cells_noalloc: bad argument");
    const bool bn = c->desc.pde == ADERDG_PDE_BN;
    if (!bn && c->desc.pde != ADERDG_PDE_EULER)
        return fail(ADERDG_EINVAL, "aderdg_predictor_cells_noalloc: Euler or Baer-Nunziato contexts only");
    if (bn && !lShi) return fail(ADERDG_EINVAL, "aderdg_predictor_cells_noalloc: lShi is required for Baer-Nunziato");
    const size_t n = c->n, V = c->V, nc = (size_t)ncells;
    // uh, qhi, Shi (V n^2 each), Fhi (2 V n^2), qbnd and Fbnd (4 V n each) per cell
    const size_t per_cell = V * n * n * 5 + 8 * V * n;
    const size_t max_cells = kScratchDoubles / per_cell;
    if (nc > max_cells)
        return fail(ADERDG_ECAPACITY, "aderdg_predictor_cells_noalloc: %d cells exceed the %zu that the %zu-byte "
                    "staging array holds for nVar %zu, N %d", ncells, max_cells, kScratchDoubles * sizeof(double), V,
                    c->N);
    const int dev = c->desc.device;
    double *base = (dev >= 0 && dev < kMaxScratchDevices) ? g_scratch[dev] : nullptr;
    if (!base) return fail(ADERDG_ESTATE, "aderdg_predictor_cells_noalloc: staging array of device %d not resolved", dev);
    if (hipSetDevice(dev) != hipSuccess) return fail(ADERDG_EHIP, "hipSetDevice(%d) failed", dev);
    double *d_uh = base, *d_q = d_uh + V * n * n * nc, *d_f = d_q + V * n * n * nc, *d_qb = d_f + 2 * V * n * n * nc;
    double *d_fb = d_qb + 4 * V * n * nc, *d_s = d_fb + 4 * V * n * nc;
    // every output element of the ncells cells is written by the kernel: nothing needs clearing
    hipError_t e = hipMemcpyAsync(d_uh, luh, V * n * n * nc * sizeof(double), hipMemcpyHostToDevice, c->stream);
    MeshArgs m = c->g.mesh(c->slot);
    m.ncell = ncells;
    TestBufs t{d_uh, d_q, d_f, d_qb, d_fb, d_s};
    if (e == hipSuccess) e = launch_test_predictor(c->N, c->g.variant, m, c->g.b, t, dt, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(lqhi, d_q, V * n * n * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(lFhi, d_f, 2 * V * n * n * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(lqbnd, d_qb, 4 * V * n * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(lFbnd, d_fb, 4 * V * n * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && bn)
        e = hipMemcpyAsync(lShi, d_s, V * n * n * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(ADERDG_EHIP, "aderdg_predictor_cells_noalloc: %s", hipGetErrorString(e));
    return 0;
}

int aderdg_predictor_cells_noalloc_max(int pde, int N)
{
    if ((pde != ADERDG_PDE_EULER && pde != ADERDG_PDE_BN) || N < 1 || N > kMaxDegree) return 0;
    const size_t n = N + 1, V = pde == ADERDG_PDE_BN ? 9 : 4;
    return (int)(kScratchDoubles / (V * n * n * 5 + 8 * V * n));
}

namespace {
// Device arrays of one per-cell call, freed when the call returns.
struct CellArrays {
    std::vector<void *> held;
    ~CellArrays()
    {
        for (void *p : held) (void)hipFree(p);
    }
    void *bytes(size_t count)
    {
        void *p = nullptr;
        if (hipMalloc(&p, count > 0 ? count : 8) != hipSuccess) return nullptr;
        held.push_back(p);
        return p;
    }
    double *doubles(size_t count) { return static_cast<double *>(bytes(count * sizeof(double))); }
    int *ints(size_t count) { return static_cast<int *>(bytes(count * sizeof(int))); }
};

// Common checks of the per-cell entries: an Euler or Baer-Nunziato context and ncells >= 1.
int cells_check(aderdg_ctx *c, int ncells, const char *who)
{
    if (!c || ncells < 1) return fail(ADERDG_EINVAL, "%s: bad argument", who);
    if (c->desc.pde != ADERDG_PDE_EULER && c->desc.pde != ADERDG_PDE_BN)
        return fail(ADERDG_EINVAL, "%s: Euler or Baer-Nunziato contexts only", who);
    if (hipSetDevice(c->desc.device) != hipSuccess) return fail(ADERDG_EHIP, "%s: hipSetDevice failed", who);
    return 0;
}

hipError_t up(double *d, const double *h, size_t count, hipStream_t st)
{
    return hipMemcpyAsync(d, h, count * sizeof(double), hipMemcpyHostToDevice, st);
}

// Launch kind on the uploaded arrays, read back the outputs (host, device, doubles) and wait.
int cells_run(aderdg_ctx *c, int kind, int ncells, const CellBufs &d, double dt, double argmax, hipError_t e,
              std::initializer_list<std::tuple<void *, const void *, size_t>> outs, const char *who)
{
    const MeshArgs m = c->g.mesh(c->slot);
    if (e == hipSuccess) e = launch_cells(kind, m, c->g.b.op, c->N, ncells, d, dt, argmax, c->stream);
    for (const auto &o : outs)
        if (e == hipSuccess)
            e = hipMemcpyAsync(std::get<0>(o), std::get<1>(o), std::get<2>(o), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(ADERDG_EHIP, "%s: %s", who, hipGetErrorString(e));
    return 0;
}
}  // namespace

int aderdg_volume_integral_cells(aderdg_ctx *c, int ncells, double *lduh, const double *lFhi, const double *lShi)
{
    const char *who = "aderdg_volume_integral_cells";
    if (int r = cells_check(c, ncells, who)) return r;
    const bool bn = c->desc.pde == ADERDG_PDE_BN;
    if (!lduh || !lFhi || (bn && !lShi)) return fail(ADERDG_EINVAL, "%s: null array%s", who, bn ? " (lShi is required)" : "");
    const size_t blk = (size_t)c->V * c->n * c->n * ncells;
    CellArrays a;
    CellBufs d;
    d.lduh = a.doubles(blk);
    double *fhi = a.doubles(2 * blk), *shi = bn ? a.doubles(blk) : nullptr;
    if (!d.lduh || !fhi || (bn && !shi)) return fail(ADERDG_EHIP, "%s: device allocation failed", who);
    d.fhi = fhi;
    d.shi = shi;
    hipError_t e = up(fhi, lFhi, 2 * blk, c->stream);
    if (e == hipSuccess && bn) e = up(shi, lShi, blk, c->stream);
    return cells_run(c, kCellVolume, ncells, d, 0.0, 0.0, e, {{lduh, d.lduh, blk * sizeof(double)}}, who);
}

int aderdg_surface_integral_cells(aderdg_ctx *c, int ncells, double *lduh, const double *lFl, const double *lFr,
                                  const double *lGb, const double *lGt)
{
    const char *who = "aderdg_surface_integral_cells";
    if (int r = cells_check(c, ncells, who)) return r;
    if (!lduh || !lFl || !lFr || !lGb || !lGt) return fail(ADERDG_EINVAL, "%s: null array", who);
    const size_t blk = (size_t)c->V * c->n * c->n * ncells, fln = (size_t)c->V * c->n * ncells;
    CellArrays a;
    CellBufs d;
    d.lduh = a.doubles(blk);
    double *f[4];
    for (double *&p : f) p = a.doubles(fln);
    if (!d.lduh || !f[0] || !f[1] || !f[2] || !f[3]) return fail(ADERDG_EHIP, "%s: device allocation failed", who);
    d.fl = f[0];
    d.fr = f[1];
    d.gb = f[2];
    d.gt = f[3];
    hipError_t e = up(d.lduh, lduh, blk, c->stream);
    const double *h[4] = {lFl, lFr, lGb, lGt};
    for (int k = 0; k < 4 && e == hipSuccess; k++) e = up(f[k], h[k], fln, c->stream);
    return cells_run(c, kCellSurface, ncells, d, 0.0, 0.0, e, {{lduh, d.lduh, blk * sizeof(double)}}, who);
}

int aderdg_element_update_cells(aderdg_ctx *c, int ncells, double dt, double *luh, double *lduh)
{
    const char *who = "aderdg_element_update_cells";
    if (int r = cells_check(c, ncells, who)) return r;
    if (!luh || !lduh) return fail(ADERDG_EINVAL, "%s: null array", who);
    const size_t blk = (size_t)c->V * c->n * c->n * ncells;
    CellArrays a;
    CellBufs d;
    d.luh = a.doubles(blk);
    d.lduh = a.doubles(blk);
    if (!d.luh || !d.lduh) return fail(ADERDG_EHIP, "%s: device allocation failed", who);
    hipError_t e = up(d.luh, luh, blk, c->stream);
    if (e == hipSuccess) e = up(d.lduh, lduh, blk, c->stream);
    return cells_run(c, kCellUpdate, ncells, d, dt, 0.0, e,
                     {{luh, d.luh, blk * sizeof(double)}, {lduh, d.lduh, blk * sizeof(double)}}, who);
}

int aderdg_dmp_cells(aderdg_ctx *c, int ncells, int32_t *dmpresult, const double *arguh, const double *lmin,
                     const double *lmax, double argmax)
{
    const char *who = "aderdg_dmp_cells";
    if (int r = cells_check(c, ncells, who)) return r;
    if (!dmpresult || !arguh || !lmin || !lmax) return fail(ADERDG_EINVAL, "%s: null array", who);
    const size_t V = c->V, n = c->n, s = c->s, nc = ncells, blk = V * n * n * nc;
    CellArrays a;
    CellBufs d;
    d.luh = a.doubles(blk);
    double *lo = a.doubles(V * nc), *hi = a.doubles(V * nc);
    d.work = a.doubles(V * (s * s + n * n) * nc);
    d.result = a.ints(nc);
    if (!d.luh || !lo || !hi || !d.work || !d.result) return fail(ADERDG_EHIP, "%s: device allocation failed", who);
    d.lmin = lo;
    d.lmax = hi;
    static_assert(sizeof(int) == sizeof(int32_t), "dmpresult is read back as int");
    hipError_t e = up(d.luh, arguh, blk, c->stream);
    if (e == hipSuccess) e = up(lo, lmin, V * nc, c->stream);
    if (e == hipSuccess) e = up(hi, lmax, V * nc, c->stream);
    return cells_run(c, kCellDmp, ncells, d, 0.0, argmax, e, {{dmpresult, d.result, nc * sizeof(int)}}, who);
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

namespace {
// Advection-diffusion predictor + volume integral (Advection-Diffusion/space_time_predictor.f90:10-155,
// Advection-Diffusion/volume_integral.f90); `fused` adds the convective part of the predictor flux.
double ad_flops_predictor(int N, bool fused)
{
    const double n = N + 1;
    const double per_slice = 2 * n * n * (2 * n + 2) + n * n * (fused ? 6 : 2) + 2 * n * (3 + n * (2 * n + 2));
    const double picard = n * (n * per_slice + n * n * (2 + n * (2 * n + 1)));
    return 3 * n * n * n + picard + 5 * n * n * 2 * n + 2 * n * 6 * 2 * n + 2 * n * n * (2 * n + 3);
}
}  // namespace

double aderdg_flops_predictor_ad(int N, int fused)
{
    return N >= 1 ? ad_flops_predictor(N, fused != 0) : 0.0;
}

// Operation counts per cell-step with the oracle's convention (+ - * / sqrt = 1).
// Predictor launch: Euler/space_time_predictor.f90 + Euler/volume_integral.f90
// (Baer-Nunziato: + gradients, NCP source and its time average).
double aderdg_flops_predictor(int pde, int N)
{
    if (pde == ADERDG_PDE_AD && N >= 1) return ad_flops_predictor(N, false);
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
