// batch.hip -- the batched engine of include/aderdg.h (aderdg_batch_*): one
// solve group per (pde, N), a single captured graph that steps every group,
// admission into the groups' arenas and retirement of finished solves.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "../../include/aderdg.h"
#include "batch.hpp"
#include "capi_common.hpp"

using namespace aderdg;

namespace {

// Health of one solve: out[0] += non-finite DOFs of uh, out[1] += troubled cells
// (recompute 1 of the last step).  Grid-stride over the solve's DOFs and cells.
__global__ void k_solve_health(const double *__restrict__ uh, const int *__restrict__ rc_final, long long ndof,
                               int ncell, unsigned long long *out)
{
    unsigned long long bad = 0, troubled = 0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < ndof; k += stride)
        bad += __builtin_isfinite(uh[k]) ? 0ull : 1ull;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < ncell; k += stride)
        troubled += rc_final[k] == 1 ? 1ull : 0ull;
    for (int o = 32; o > 0; o >>= 1) {
        bad += __shfl_down(bad, o);
        troubled += __shfl_down(troubled, o);
    }
    if ((threadIdx.x & 63) == 0) {
        if (bad) atomicAdd(out, bad);
        if (troubled) atomicAdd(out + 1, troubled);
    }
}

}  // namespace

namespace aderdg {

int Batch::init(int device_, int persist_)
{
    device = device_;
    HIPCHK(hipSetDevice(device));
    int cus = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    persist = persist_ > 0 ? persist_ : 8 * std::max(cus, 1);
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&copy, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&upload_done, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&removed, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&states_ready, hipEventDisableTiming));
    return 0;
}

void Batch::release()
{
    if (stream) (void)hipStreamSynchronize(stream);
    if (copy) (void)hipStreamSynchronize(copy);
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
    for (auto &g : groups) g->release();
    groups.clear();
    for (auto &t : pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    pending.clear();
    for (auto e : free_events) (void)hipEventDestroy(e);
    free_events.clear();
    for (auto p : pinned_states)
        if (p) (void)hipHostFree(p);
    pinned_states.clear();
    if (health_buf) (void)hipFree(health_buf);
    health_buf = nullptr;
    for (auto e : join) (void)hipEventDestroy(e);
    join.clear();
    if (fork) (void)hipEventDestroy(fork);
    fork = nullptr;
    for (auto b : branch) (void)hipStreamDestroy(b);
    branch.clear();
    hipEvent_t evs[] = {upload_done, removed, states_ready};
    for (auto e : evs)
        if (e) (void)hipEventDestroy(e);
    upload_done = removed = states_ready = nullptr;
    if (copy) (void)hipStreamDestroy(copy);
    if (stream) (void)hipStreamDestroy(stream);
    copy = stream = nullptr;
}

int Batch::group_of(int pde, int N) const
{
    for (size_t k = 0; k < groups.size(); k++)
        if (groups[k]->pde == pde && groups[k]->N == N) return (int)k;
    return -1;
}

int64_t Batch::footprint(int pde, int N, int imax, int jmax, double lh_fraction)
{
    int64_t pc, px, py, ps, pl;
    Group::footprint(pde, N, &pc, &px, &py, &ps, &pl);
    const int64_t nc = (int64_t)imax * jmax;
    return pc * nc + px * (int64_t)(imax + 1) * jmax + py * (int64_t)imax * (jmax + 1) + ps +
           pl * Group::lh_slots(pde, nc, lh_fraction);
}

double Batch::lh_fraction(int pde, int N) const
{
    const int g = group_of(pde, N);
    return g < 0 ? 1.0 : groups[g]->caps.lh_fraction;
}

int Batch::reserve(int pde, int N, int64_t max_cells, int max_solves, double lh_fraction)
{
    if (pde != ADERDG_PDE_EULER && pde != ADERDG_PDE_BN && pde != ADERDG_PDE_AD)
        return fail(ADERDG_EINVAL, "aderdg_batch_reserve: bad pde %d", pde);
    if (N < 1 || N > kMaxDegree) return fail(ADERDG_EINVAL, "aderdg_batch_reserve: N=%d outside 1..%d", N, kMaxDegree);
    if (max_cells < 1 || max_cells > (1LL << 28) || max_solves < 1 || max_solves > (1 << 20))
        return fail(ADERDG_EINVAL, "aderdg_batch_reserve: bad capacity (%lld cells, %d solves)", (long long)max_cells,
                    max_solves);
    if (!(lh_fraction > 0.0 && lh_fraction <= 1.0))
        return fail(ADERDG_EINVAL, "aderdg_batch_reserve: limited-cell fraction %g outside (0, 1]", lh_fraction);
    if (group_of(pde, N) >= 0) return fail(ADERDG_EINVAL, "aderdg_batch_reserve: (pde %d, N %d) already reserved", pde, N);
    Operators ops;
    if (!build_operators(N, ops)) return fail(ADERDG_ESINGULAR, "aderdg_batch_reserve: operator construction failed");
    // faces: (IMAX+1)*JMAX = cells + JMAX <= cells * 9/8 when IMAX >= 8, and alike for y
    const int64_t faces = max_cells + max_cells / 8 + 1;
    const GroupCaps caps{max_cells, faces, faces, max_solves, lh_fraction};
    auto g = std::make_unique<Group>();
    int r = g->init(pde, N, caps, ops, persist, stream);
    if (r) {
        g->release();
        return fail(r == ADERDG_EINVAL ? r : ADERDG_EHIP, "aderdg_batch_reserve: arena allocation failed");
    }
    SolveState *ps = nullptr;
    if (hipHostMalloc((void **)&ps, sizeof(SolveState) * max_solves, hipHostMallocDefault) != hipSuccess) {
        g->release();
        return fail(ADERDG_EHIP, "aderdg_batch_reserve: hipHostMalloc failed");
    }
    int64_t pc, px, py, psl, pl;
    Group::footprint(pde, N, &pc, &px, &py, &psl, &pl);
    reserved += pc * max_cells + (px + py) * faces + psl * max_solves + pl * Group::lh_pool(pde, max_cells, max_solves, lh_fraction);
    groups.push_back(std::move(g));
    pinned_states.push_back(ps);
    ms.emplace_back(ADERDG_TIMING_SLOTS, 0.0);
    launches.emplace_back(ADERDG_TIMING_SLOTS, 0);
    graph_stale = true;
    return 0;
}

int Batch::add(const aderdg_desc *d, const double *uh, const aderdg_ic *ic, double time, double tend, int max_steps,
               int stop_at_tend, int *id)
{
    int r = check_desc(d, "aderdg_batch_add");
    if (r) return r;
    if ((!uh && !ic) || !id || max_steps < 1) return fail(ADERDG_EINVAL, "aderdg_batch_add: bad argument");
    if (ic && !ic_fits(d->pde, ic->kind))
        return fail(ADERDG_EINVAL, "aderdg_batch_add_ic: initial condition kind %d does not fit pde %d", ic->kind,
                    d->pde);
    const aderdg_ops &o = d->ops;
    if (o.wGPN || o.Kxi || o.iK1 || o.F0 || o.FLcoeff || o.FRcoeff || o.uh2lim || o.lim2uh || o.uh2lob)
        return fail(ADERDG_EINVAL, "aderdg_batch_add: batched solves use the built-in operators (ops must be NULL)");
    if (d->device != device) return fail(ADERDG_EINVAL, "aderdg_batch_add: solve on device %d, batch on %d", d->device, device);
    const int gi = group_of(d->pde, d->N);
    if (gi < 0) return fail(ADERDG_EINVAL, "aderdg_batch_add: no arena reserved for (pde %d, N %d)", d->pde, d->N);
    Group &g = *groups[gi];
    SolveState st{};
    st.time = time;
    st.tend = tend;
    st.dtmin = ~0ull;
    st.active = 1;
    st.seed = 1;
    st.max_steps = max_steps;
    st.stop_at_tend = stop_at_tend ? 1 : 0;
    int slot = -1;
    r = g.add(mesh_from_desc(d), st, &slot);
    if (r == ADERDG_EFULL) return fail(r, "aderdg_batch_add: arena (pde %d, N %d) full", d->pde, d->N);
    if (r) return fail(r, "aderdg_batch_add: placement failed");
    hipError_t e = hipSuccess;
    if (uh) {
        // the copy may reuse cells of a removed solve: order it after the removal
        HIPCHK(hipStreamWaitEvent(copy, removed, 0));
        if ((r = g.upload_uh(slot, uh, 0, copy))) return fail(r, "aderdg_batch_add: upload failed");
        HIPCHK(hipEventRecord(upload_done, copy));
        HIPCHK(hipStreamWaitEvent(stream, upload_done, 0));
    } else {
        IcSpec s{};
        s.kind = ic->kind;
        for (int k = 0; k < kIcParams; k++) s.p[k] = ic->p[k];
        e = launch_init(g.pde, g.N, g.b, s, slot, d->imax * d->jmax, 0, stream);
    }
    if (e == hipSuccess) e = g.launch(kPartSeed);
    if (e == hipSuccess) e = g.launch(kPartSeedDone);
    if (e != hipSuccess) return fail(ADERDG_EHIP, "aderdg_batch_add: seed launch: %s", hipGetErrorString(e));
    const int h = next_id++;
    solves[h] = Handle{gi, slot};
    *id = h;
    return 0;
}

int Batch::build_graph()
{
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
    hipGraph_t gr = nullptr;
    // arenas are independent: one graph branch each, so one arena's compute-bound predictor
    // can run beside another's memory-bound face and update passes and fill its tails
    while (branch.size() + 1 < groups.size()) {
        hipStream_t b = nullptr;
        hipEvent_t j = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&b, hipStreamNonBlocking));
        branch.push_back(b);
        HIPCHK(hipEventCreateWithFlags(&j, hipEventDisableTiming));
        join.push_back(j);
    }
    if (!fork) HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    static const int parts[] = {kPartRanges, kPartPredictor, kPartFace, kPartUpdate, kPartSpread, kPartSubcell,
                                kPartAdvance};
    if (groups.size() > 1) e = hipEventRecord(fork, stream);
    for (size_t k = 1; k < groups.size() && e == hipSuccess; k++) e = hipStreamWaitEvent(branch[k - 1], fork, 0);
    for (size_t k = 0; k < groups.size(); k++) {
        const Group &g = *groups[k];
        hipStream_t s = k == 0 ? stream : branch[k - 1];
        if (e == hipSuccess) e = launch_dt(g.b, s, 0.0, 0.0, 0, g.dt_grid());
        for (int p : parts)
            if (e == hipSuccess) e = g.launch_on(p, s, groups.size() > 1);
    }
    for (size_t k = 1; k < groups.size() && e == hipSuccess; k++) {
        e = hipEventRecord(join[k - 1], branch[k - 1]);
        if (e == hipSuccess) e = hipStreamWaitEvent(stream, join[k - 1], 0);
    }
    hipError_t e2 = hipStreamEndCapture(stream, &gr);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (gr) (void)hipGraphDestroy(gr);
        return fail(ADERDG_EHIP, "aderdg_batch: step capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    e = hipGraphInstantiate(&graph, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) return fail(ADERDG_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    graph_stale = false;
    return 0;
}

int Batch::collect_timing()
{
    HIPCHK(hipStreamSynchronize(stream));
    for (auto &t : pending) {
        float m = 0.f;
        if (hipEventElapsedTime(&m, t.a, t.b) == hipSuccess) {
            ms[t.group][t.slot] += m;
            launches[t.group][t.slot] += 1;
        }
        free_events.push_back(t.a);
        free_events.push_back(t.b);
    }
    pending.clear();
    return 0;
}

int Batch::step(int nsteps)
{
    if (nsteps < 0) return fail(ADERDG_EINVAL, "aderdg_batch_step: negative step count");
    if (groups.empty()) return fail(ADERDG_ESTATE, "aderdg_batch_step: nothing reserved");
    if (timing) {  // the graph's launches one by one, each between two events (each alone on the
                   // GPU, so the matrix-core predictor takes its full grid, not the shared one)
        // timing slots of include/aderdg.h: dt, predictor, update, spread, subcell, face, ranges; -1 untimed
        static const int parts[] = {kPartDt, kPartRanges, kPartPredictor, kPartFace, kPartUpdate, kPartSpread,
                                    kPartSubcell, kPartAdvance};
        static const int slots[] = {0, 6, 1, 5, 2, 3, 4, -1};
        constexpr int nparts = sizeof parts / sizeof parts[0];
        for (int k = 0; k < nsteps; k++)
            for (size_t gi = 0; gi < groups.size(); gi++)
                for (int p = 0; p < nparts; p++) {
                    hipEvent_t e0 = nullptr, e1 = nullptr;
                    for (hipEvent_t *e : {&e0, &e1}) {
                        if (!free_events.empty()) {
                            *e = free_events.back();
                            free_events.pop_back();
                        } else {
                            HIPCHK(hipEventCreate(e));
                        }
                    }
                    HIPCHK(hipEventRecord(e0, stream));
                    hipError_t e = parts[p] == kPartDt ? launch_dt(groups[gi]->b, stream, 0.0, 0.0, 0, groups[gi]->dt_grid())
                                                       : groups[gi]->launch(parts[p]);
                    if (e != hipSuccess) return fail(ADERDG_EHIP, "aderdg_batch_step: launch: %s", hipGetErrorString(e));
                    HIPCHK(hipEventRecord(e1, stream));
                    if (slots[p] >= 0)
                        pending.push_back({(int)gi, slots[p], e0, e1});
                    else {
                        free_events.push_back(e0);
                        free_events.push_back(e1);
                    }
                }
        return 0;
    }
    if (graph_stale) {
        int r = build_graph();
        if (r) return r;
    }
    for (int k = 0; k < nsteps; k++) HIPCHK(hipGraphLaunch(graph, stream));
    return 0;
}

int Batch::snapshot_async()
{
    for (size_t k = 0; k < groups.size(); k++)
        if (groups[k]->read_states_async(pinned_states[k])) return fail(ADERDG_EHIP, "aderdg_batch: state readback failed");
    HIPCHK(hipEventRecord(states_ready, stream));
    return 0;
}

bool Batch::snapshot_ready() const { return hipEventQuery(states_ready) == hipSuccess; }

int Batch::snapshot()
{
    int r = snapshot_async();
    if (r) return r;
    HIPCHK(hipEventSynchronize(states_ready));
    return 0;
}

const SolveState *Batch::state_of(int id) const
{
    auto it = solves.find(id);
    if (it == solves.end()) return nullptr;
    return &pinned_states[it->second.group][it->second.slot];
}

int Batch::download(int id, double *uh)
{
    auto it = solves.find(id);
    if (it == solves.end()) return fail(ADERDG_EINVAL, "aderdg_batch_download: unknown id %d", id);
    int r = snapshot();
    if (r) return r;
    Group &g = *groups[it->second.group];
    const int par = pinned_states[it->second.group][it->second.slot].parity;
    if ((r = g.download_uh(it->second.slot, uh, par))) return fail(r, "aderdg_batch_download: copy failed");
    HIPCHK(hipStreamSynchronize(stream));
    return 0;
}

int Batch::limiter(int id, int32_t *status, int32_t *recompute, double *Lh, double *lmin, double *lmax)
{
    auto it = solves.find(id);
    if (it == solves.end()) return fail(ADERDG_EINVAL, "aderdg_batch_limiter: unknown id %d", id);
    int r = snapshot();
    if (r) return r;
    Group &g = *groups[it->second.group];
    const int sl = it->second.slot;
    const int par = pinned_states[it->second.group][sl].parity;
    const size_t c0 = g.cell0(sl), nc = g.mesh(sl).ncell, V = g.V, G = g.b.G;
    const size_t blk = (size_t)g.V * g.s * g.s;
    const SolveArgs &sa = g.args(sl);
    std::vector<int32_t> st(nc), slot;
    std::vector<double> pool, mn, mx;
    HIPCHK(hipMemcpyAsync(st.data(), g.b.status[par] + c0, nc * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    if (recompute) HIPCHK(hipMemcpyAsync(recompute, g.b.rc_final + c0, nc * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    if (Lh && sa.lhcap > 0) {  // the solve's pool segment and the slots of its limited cells
        slot.resize(nc);
        pool.resize((size_t)sa.lhcap * blk);
        HIPCHK(hipMemcpyAsync(slot.data(), g.b.lh_slot[par] + c0, nc * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(pool.data(), g.b.lh[par] + (size_t)sa.lh0 * blk, pool.size() * sizeof(double),
                              hipMemcpyDeviceToHost, stream));
    }
    // lmin/lmax are [v][cell] planes over the arena
    for (size_t v = 0; v < V && (lmin || lmax); v++) {
        if (lmin) {
            mn.resize(V * nc);
            HIPCHK(hipMemcpyAsync(mn.data() + v * nc, g.b.lmin + v * G + c0, nc * sizeof(double), hipMemcpyDeviceToHost,
                                  stream));
        }
        if (lmax) {
            mx.resize(V * nc);
            HIPCHK(hipMemcpyAsync(mx.data() + v * nc, g.b.lmax + v * G + c0, nc * sizeof(double), hipMemcpyDeviceToHost,
                                  stream));
        }
    }
    HIPCHK(hipStreamSynchronize(stream));
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

int Batch::error(int id, int *err)
{
    auto it = solves.find(id);
    if (it == solves.end()) return fail(ADERDG_EINVAL, "aderdg_batch_error: unknown id %d", id);
    int r = snapshot();
    if (r) return r;
    *err = pinned_states[it->second.group][it->second.slot].failed ? ADERDG_ECAPACITY : 0;
    return 0;
}

int Batch::health(int id, int64_t *nonfinite, int64_t *troubled)
{
    auto it = solves.find(id);
    if (it == solves.end()) return fail(ADERDG_EINVAL, "aderdg_batch_health: unknown id %d", id);
    int r = snapshot();
    if (r) return r;
    if (!health_buf) HIPCHK(hipMalloc((void **)&health_buf, 2 * sizeof(unsigned long long)));
    Group &g = *groups[it->second.group];
    const int sl = it->second.slot;
    const int par = pinned_states[it->second.group][sl].parity;
    const size_t c0 = g.cell0(sl), nc = g.mesh(sl).ncell, blk = (size_t)g.V * g.n * g.n;
    HIPCHK(hipMemsetAsync(health_buf, 0, 2 * sizeof(unsigned long long), stream));
    const long long ndof = (long long)(nc * blk);
    const int blocks = (int)std::min<long long>(1024, (ndof + 255) / 256);
    hipLaunchKernelGGL(k_solve_health, dim3(blocks), dim3(256), 0, stream, g.b.uh[par] + c0 * blk,
                       g.b.rc_final + c0, ndof, (int)nc, health_buf);
    HIPCHK(hipGetLastError());
    unsigned long long h[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(h, health_buf, sizeof(h), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (nonfinite) *nonfinite = (int64_t)h[0];
    if (troubled) *troubled = (int64_t)h[1];
    return 0;
}

int Batch::remove(int id)
{
    auto it = solves.find(id);
    if (it == solves.end()) return fail(ADERDG_EINVAL, "aderdg_batch_remove: unknown id %d", id);
    int r = groups[it->second.group]->remove(it->second.slot);
    if (r) return fail(r, "aderdg_batch_remove: failed");
    HIPCHK(hipEventRecord(removed, stream));
    solves.erase(it);
    return 0;
}

}  // namespace aderdg

struct aderdg_batch {
    Batch eng;
};

extern "C" {

int aderdg_batch_create(const aderdg_batch_desc *d, aderdg_batch **out)
{
    if (!d || !out) return fail(ADERDG_EINVAL, "aderdg_batch_create: null argument");
    *out = nullptr;
    auto *b = new aderdg_batch();
    int r = b->eng.init(d->device, d->persist_blocks);
    if (r) {
        b->eng.release();
        delete b;
        return r;
    }
    *out = b;
    return 0;
}

int aderdg_batch_destroy(aderdg_batch *b)
{
    if (!b) return 0;
    b->eng.release();
    delete b;
    return 0;
}

int aderdg_batch_reserve(aderdg_batch *b, int pde, int N, int64_t max_cells, int max_solves)
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_reserve: null batch");
    return b->eng.reserve(pde, N, max_cells, max_solves);
}

int aderdg_batch_reserve_limited(aderdg_batch *b, int pde, int N, int64_t max_cells, int max_solves,
                                 double limited_fraction)
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_reserve_limited: null batch");
    return b->eng.reserve(pde, N, max_cells, max_solves, limited_fraction);
}

int64_t aderdg_solve_footprint_limited(int pde, int N, int imax, int jmax, double limited_fraction)
{
    if ((pde != ADERDG_PDE_EULER && pde != ADERDG_PDE_BN && pde != ADERDG_PDE_AD) || N < 1 || N > kMaxDegree ||
        imax < 1 || jmax < 1 || !(limited_fraction > 0.0 && limited_fraction <= 1.0))
        return -1;
    return Batch::footprint(pde, N, imax, jmax, limited_fraction);
}

int64_t aderdg_solve_footprint(int pde, int N, int imax, int jmax)
{
    return aderdg_solve_footprint_limited(pde, N, imax, jmax, 1.0);
}

int aderdg_batch_error(aderdg_batch *b, int id, int *error)
{
    if (!b || !error) return fail(ADERDG_EINVAL, "aderdg_batch_error: null argument");
    return b->eng.error(id, error);
}

int64_t aderdg_batch_reserved_bytes(aderdg_batch *b) { return b ? b->eng.reserved : -1; }

int aderdg_batch_add(aderdg_batch *b, const aderdg_desc *solve, const double *uh, double time, double tend,
                     int max_steps, int stop_at_tend, int *id)
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_add: null batch");
    return b->eng.add(solve, uh, nullptr, time, tend, max_steps, stop_at_tend, id);
}

int aderdg_batch_add_ic(aderdg_batch *b, const aderdg_desc *solve, const aderdg_ic *ic, double time, double tend,
                        int max_steps, int stop_at_tend, int *id)
{
    if (!b || !ic) return fail(ADERDG_EINVAL, "aderdg_batch_add_ic: null argument");
    return b->eng.add(solve, nullptr, ic, time, tend, max_steps, stop_at_tend, id);
}

int aderdg_batch_step(aderdg_batch *b, int nsteps)
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_step: null batch");
    return b->eng.step(nsteps);
}

int aderdg_batch_sync(aderdg_batch *b)
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_sync: null batch");
    HIPCHK(hipStreamSynchronize(b->eng.copy));
    HIPCHK(hipStreamSynchronize(b->eng.stream));
    return 0;
}

int aderdg_batch_query(aderdg_batch *b, int id, double *time, double *dt, int *steps, int *active)
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_query: null batch");
    if (!b->eng.state_of(id)) return fail(ADERDG_EINVAL, "aderdg_batch_query: unknown id %d", id);
    int r = b->eng.snapshot();
    if (r) return r;
    const SolveState *s = b->eng.state_of(id);
    if (time) *time = s->time;
    if (dt) *dt = s->dt;
    if (steps) *steps = s->steps;
    if (active) *active = s->active;
    return 0;
}

int aderdg_batch_finished(aderdg_batch *b, int *ids, int max_ids, int *count)
{
    if (!b || !count || (max_ids > 0 && !ids)) return fail(ADERDG_EINVAL, "aderdg_batch_finished: bad argument");
    int r = b->eng.snapshot();
    if (r) return r;
    int c = 0;
    for (const auto &kv : b->eng.solves) {
        const SolveState *s = b->eng.state_of(kv.first);
        if (!s->active && c < max_ids) ids[c++] = kv.first;
    }
    *count = c;
    return 0;
}

int aderdg_batch_download(aderdg_batch *b, int id, double *uh)
{
    if (!b || !uh) return fail(ADERDG_EINVAL, "aderdg_batch_download: null argument");
    return b->eng.download(id, uh);
}

int aderdg_batch_limiter(aderdg_batch *b, int id, int32_t *status, int32_t *recompute)
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_limiter: null batch");
    return b->eng.limiter(id, status, recompute);
}

int aderdg_batch_limiter_state(aderdg_batch *b, int id, int32_t *status, int32_t *recompute, double *Lh,
                               double *lmin, double *lmax)
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_limiter_state: null batch");
    return b->eng.limiter(id, status, recompute, Lh, lmin, lmax);
}

int aderdg_batch_health(aderdg_batch *b, int id, int64_t *nonfinite, int64_t *troubled)
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_health: null batch");
    return b->eng.health(id, nonfinite, troubled);
}

int aderdg_batch_remove(aderdg_batch *b, int id)
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_remove: null batch");
    return b->eng.remove(id);
}

int aderdg_batch_timing_enable(aderdg_batch *b, int on)
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_timing_enable: null batch");
    int r = b->eng.collect_timing();
    if (r) return r;
    b->eng.timing = on != 0;
    for (auto &v : b->eng.ms) std::fill(v.begin(), v.end(), 0.0);
    for (auto &v : b->eng.launches) std::fill(v.begin(), v.end(), 0);
    return 0;
}

int aderdg_batch_timing_get(aderdg_batch *b, int pde, int N, double ms[ADERDG_TIMING_SLOTS],
                            int64_t launches[ADERDG_TIMING_SLOTS])
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_timing_get: null batch");
    const int gi = b->eng.group_of(pde, N);
    if (gi < 0) return fail(ADERDG_EINVAL, "aderdg_batch_timing_get: no arena (pde %d, N %d)", pde, N);
    int r = b->eng.collect_timing();
    if (r) return r;
    for (int k = 0; k < ADERDG_TIMING_SLOTS; k++) {
        if (ms) ms[k] = b->eng.ms[gi][k];
        if (launches) launches[k] = b->eng.launches[gi][k];
    }
    return 0;
}

int aderdg_batch_usage(aderdg_batch *b, int *solves, int64_t *cells)
{
    if (!b) return fail(ADERDG_EINVAL, "aderdg_batch_usage: null batch");
    int64_t c = 0;
    for (auto &g : b->eng.groups) c += g->cells_in_use();
    if (solves) *solves = (int)b->eng.solves.size();
    if (cells) *cells = c;
    return 0;
}

}  // extern "C"
