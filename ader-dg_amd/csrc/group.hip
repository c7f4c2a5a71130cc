// group.hip -- host side of a solve group (see group.hpp).
#include "group.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/aderdg.h"

namespace aderdg {

int Staging::init(size_t bytes)
{
    release();
    if (hipHostMalloc((void **)&buf_, bytes, hipHostMallocDefault) != hipSuccess) return ADERDG_EHIP;
    cap_ = bytes;
    pos_ = 0;
    return 0;
}

void Staging::release()
{
    for (auto &p : pending_) {
        (void)hipEventSynchronize(p.ev);
        (void)hipEventDestroy(p.ev);
    }
    pending_.clear();
    for (auto e : free_) (void)hipEventDestroy(e);
    free_.clear();
    if (buf_) (void)hipHostFree(buf_);
    buf_ = nullptr;
    cap_ = pos_ = 0;
}

hipError_t Staging::put(void *dst, const void *src, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    if (n > cap_) return hipErrorInvalidValue;
    const size_t len = (n + 255) & ~(size_t)255;
    if (pos_ + len > cap_) pos_ = 0;
    const size_t start = pos_, end = pos_ + len;
    // wait for the chunks that overlap [start, end); retire every completed chunk
    std::vector<Pending> keep;
    for (const Pending &p : pending_) {
        bool done = hipEventQuery(p.ev) == hipSuccess;
        if (!done && p.start < end && start < p.end) {
            hipError_t e = hipEventSynchronize(p.ev);
            if (e != hipSuccess) return e;
            done = true;
        }
        if (done)
            free_.push_back(p.ev);
        else
            keep.push_back(p);
    }
    pending_.swap(keep);
    std::memcpy(buf_ + start, src, n);
    hipError_t e = hipMemcpyAsync(dst, buf_ + start, n, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    hipEvent_t ev = nullptr;
    if (!free_.empty()) {
        ev = free_.back();
        free_.pop_back();
    } else if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) {
        return e;
    }
    if ((e = hipEventRecord(ev, st)) != hipSuccess) return e;
    pending_.push_back({start, end, ev});
    pos_ = end;
    return hipSuccess;
}

void RangeAlloc::reset(int64_t cap)
{
    cap_ = cap;
    free_.clear();
    if (cap > 0) free_[0] = cap;
}

bool RangeAlloc::alloc(int64_t len, int64_t *start)
{
    if (len <= 0) {
        *start = 0;
        return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= len) {
            *start = it->first;
            const int64_t rest = it->second - len, at = it->first + len;
            free_.erase(it);
            if (rest > 0) free_[at] = rest;
            return true;
        }
    }
    return false;
}

void RangeAlloc::release(int64_t start, int64_t len)
{
    if (len <= 0) return;
    auto it = free_.emplace(start, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {  // merge with the next range
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {  // merge with the previous range
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

int64_t RangeAlloc::largest_free() const
{
    int64_t m = 0;
    for (const auto &kv : free_) m = std::max(m, kv.second);
    return m;
}

template <class T>
int Group::dalloc(T **p, size_t count)
{
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(count * sizeof(T), 8)) != hipSuccess) return ADERDG_EHIP;
    allocs_.push_back(q);
    *p = static_cast<T *>(q);
    return 0;
}

void Group::footprint(int pde, int N, int64_t *per_cell, int64_t *per_xface, int64_t *per_yface, int64_t *per_slot,
                      int64_t *per_lh)
{
    const int64_t V = pde == kPdeBn ? 9 : (pde == kPdeAd ? 1 : 4), n = N + 1, s = 2 * N + 1;
    const int64_t sides = pde == kPdeBn ? 2 : 1, ffc = pde == kPdeAd ? 2 : V;
    // uh x2, lduh, fq + ff, lmin/lmax old and new, cell_dt; ints: rc, rc_final, list, status x2,
    // lh_slot x2, cell_solve
    const int64_t dbl = 3 * V * n * n + 4 * (V + ffc) * n + 4 * V + 1;
    const int64_t ints = 8;
    // work-table entries are a few bytes per cell; counted per cell at one entry per 4 cells
    const int64_t cell = 8 * dbl + 4 * ints + 4 * 8 / 4, face = 8 * sides * V * n;
    if (per_cell) *per_cell = cell;
    if (per_xface) *per_xface = face;
    if (per_yface) *per_yface = face;
    // a slot of a batch arena also holds the alignment padding of its cell and face ranges (init)
    if (per_slot)
        *per_slot = (int64_t)(sizeof(SolveArgs) + sizeof(SolveState)) + 4 * 8 * kNumTabs + 4 +
                    (kAlign - 1) * (cell + 2 * face);
    if (per_lh) *per_lh = 2 * 8 * V * s * s;  // Lh and NewLh
}

int64_t Group::lh_slots(int pde, int64_t ncell, double fraction)
{
    if (pde == kPdeAd) return 0;  // no limiter
    if (!(fraction < 1.0)) return ncell;
    return std::min<int64_t>(ncell, (int64_t)std::ceil(fraction * (double)ncell));
}

int64_t Group::lh_pool(int pde, int64_t cells, int slots, double fraction)
{
    // every solve's share rounds up by less than one slot
    if (pde == kPdeAd) return 0;
    return std::min<int64_t>(cells, lh_slots(pde, cells, fraction) + (fraction < 1.0 ? slots : 0));
}

int Group::init(int pde_, int N_, const GroupCaps &c_in, const Operators &ops_, int persist_blocks, hipStream_t st,
                 int variant_, int upd_)
{
    // every solve's ranges start on a multiple of kAlign items: room for the padding of each slot
    // (a single context's solve starts at 0 and its planes keep the stride of its own cells)
    // and the plane strides are multiples of kAlign too, so that every plane's lines align
    GroupCaps c = c_in;
    if (c.slots > 1) {
        c.cells = align_up(c.cells + (kAlign - 1) * c.slots);
        c.xfaces = align_up(c.xfaces + (kAlign - 1) * c.slots);
        c.yfaces = align_up(c.yfaces + (kAlign - 1) * c.slots);
    }
    pde = pde_;
    N = N_;
    V = pde == kPdeBn ? 9 : (pde == kPdeAd ? 1 : 4);
    n = N + 1;
    s = 2 * N + 1;
    caps = c;
    ops = ops_;
    st_ = st;
    if (c.cells < 1 || c.slots < 1 || c.cells > (1LL << 30) || c.xfaces > (1LL << 30) || c.yfaces > (1LL << 30) ||
        !(c.lh_fraction > 0.0 && c.lh_fraction <= 1.0))
        return ADERDG_EINVAL;
    if (variant_ < 0 || upd_ < 0) {
        int tp = kPredLine, tu = kUpdShared;
        tune_kernels(pde, N, persist_blocks, ops_, st, &tp, &tu);
        variant = variant_ >= 0 ? variant_ : tp;
        upd = upd_ >= 0 ? upd_ : tu;
    } else {
        variant = variant_;
        upd = upd_;
    }
    if (!pred_variant_valid(pde, N, variant) || upd < 0 || upd >= kNumUpdVariants) return ADERDG_EINVAL;
    table_quanta(pde, N, variant, quanta);
    const int64_t items[kNumTabs] = {c.cells, c.cells, (c.xfaces + c.yfaces) * n, c.cells};
    int64_t max_entries = 0;
    // a persistent grid never exceeds the workgroups of its kernel that fit on the CUs at once
    int dev = 0;
    cus = 0;
    if (persist_blocks > 0 && (hipGetDevice(&dev) != hipSuccess ||
                               hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess))
        cus = 0;
    for (int k = 0; k < kNumTabs; k++) {
        tab_cap[k] = items[k] / quanta[k] + c.slots + 1;
        max_entries = std::max(max_entries, tab_cap[k]);
        int64_t g = persist_blocks > 0 ? std::min<int64_t>(tab_cap[k], persist_blocks) : tab_cap[k];
        if (persist_blocks > 0 && cus > 0) {
            const int r = resident_blocks_per_cu(pde, N, variant, upd, k);
            if (r > 0) g = std::min<int64_t>(g, (int64_t)r * cus);
        }
        grid[k] = (int)std::max<int64_t>(1, std::min<int64_t>(g, 1 << 30));
    }
    if (max_entries > (1LL << 30)) return ADERDG_EINVAL;
    const size_t G = c.cells, NN = (size_t)n * n, SS = (size_t)s * s, sides = pde == kPdeBn ? 2 : 1;
    const int64_t pool = lh_pool(pde, c_in.cells, c.slots, c.lh_fraction);
    b = StepBufs{};
    b.G = (int)c.cells;
    b.GX = (int)std::max<int64_t>(c.xfaces, 1);
    b.GY = (int)std::max<int64_t>(c.yfaces, 1);
    Operators *opd = nullptr;
    SolveArgs *sad = nullptr;
    int *meta = nullptr, *cs = nullptr;
    int r = 0;
    r = r ? r : dalloc(&opd, 1);
    r = r ? r : dalloc(&sad, (size_t)c.slots);
    r = r ? r : dalloc(&b.state, (size_t)c.slots);
    for (int k = 0; k < kNumTabs; k++) {
        int2 *t = nullptr;
        r = r ? r : dalloc(&t, (size_t)tab_cap[k]);
        b.tab[k] = t;
    }
    r = r ? r : dalloc(&meta, (size_t)kNumMeta);
    r = r ? r : dalloc(&prefix_dev_, (size_t)kNumTabs * (c.slots + 1));
    r = r ? r : dalloc(&cs, G);
    for (int k = 0; k < 2; k++) {
        r = r ? r : dalloc(&b.uh[k], V * NN * G);
        r = r ? r : dalloc(&b.status[k], G);
        r = r ? r : dalloc(&b.lh[k], V * SS * (size_t)pool);
        r = r ? r : dalloc(&b.lh_slot[k], G);
    }
    r = r ? r : dalloc(&b.lduh, V * NN * G);
    r = r ? r : dalloc(&b.fq, 4 * V * n * G);
    r = r ? r : dalloc(&b.ff, 4 * (size_t)(pde == kPdeAd ? 2 : V) * n * G);  // advection-diffusion: grad q
    r = r ? r : dalloc(&b.xflux, sides * V * n * (size_t)b.GX);
    r = r ? r : dalloc(&b.yflux, sides * V * n * (size_t)b.GY);
    r = r ? r : dalloc(&b.lmin_old, V * G);
    r = r ? r : dalloc(&b.lmax_old, V * G);
    r = r ? r : dalloc(&b.lmin, V * G);
    r = r ? r : dalloc(&b.lmax, V * G);
    r = r ? r : dalloc(&b.cell_dt, G);
    r = r ? r : dalloc(&b.rc, G);
    r = r ? r : dalloc(&b.rc_final, G);
    r = r ? r : dalloc(&b.list, G);
    r = r ? r : dalloc(&b.counters, (size_t)kNumCounters);
    r = r ? r : dalloc(&b.lh_count, (size_t)c.slots);
    if (r) return r;
    b.op = opd;
    b.solves = sad;
    b.meta = meta;
    b.cell_solve = cs;
    prefix_host_.assign((size_t)kNumTabs * (c.slots + 1), 0);
    if ((r = staging_.init(std::max<size_t>(1 << 20, 8 * sizeof(int) * kNumTabs * (c.slots + 1)))) != 0) return r;
    solves_.assign(c.slots, SolveArgs{});
    used_.assign(c.slots, 0);
    cell_alloc_.reset(c.cells);
    fx_alloc_.reset(c.xfaces);
    fy_alloc_.reset(c.yfaces);
    lh_alloc_.reset(pool);
    nslot_ = 0;
    cells_in_use_ = 0;
    if (hipMemcpy(opd, &ops, sizeof(Operators), hipMemcpyHostToDevice) != hipSuccess) return ADERDG_EHIP;
    if (hipMemset(b.counters, 0, kNumCounters * sizeof(int)) != hipSuccess) return ADERDG_EHIP;
    if (hipMemset(b.lh_count, 0, c.slots * sizeof(int)) != hipSuccess) return ADERDG_EHIP;
    if (hipMemset(b.state, 0, c.slots * sizeof(SolveState)) != hipSuccess) return ADERDG_EHIP;
    if (hipMemset(meta, 0, kNumMeta * sizeof(int)) != hipSuccess) return ADERDG_EHIP;
    for (int k = 0; k < 2; k++)
        if (hipMemset(b.status[k], 0, G * sizeof(int)) != hipSuccess) return ADERDG_EHIP;
    if (hipMemset(b.rc, 0, G * sizeof(int)) != hipSuccess) return ADERDG_EHIP;
    if (hipMemset(b.rc_final, 0, G * sizeof(int)) != hipSuccess) return ADERDG_EHIP;
    return 0;
}

void Group::release()
{
    if (st_) (void)hipStreamSynchronize(st_);
    staging_.release();
    for (void *p : allocs_) (void)hipFree(p);
    allocs_.clear();
    prefix_host_.clear();
    prefix_dev_ = nullptr;
    b = StepBufs{};
}

int Group::rebuild_tables()
{
    const int ns = nslot_;
    int *p = prefix_host_.data();
    for (int k = 0; k < kNumTabs; k++) {
        int *pk = p + k * (ns + 1);
        pk[0] = 0;
        for (int sl = 0; sl < ns; sl++) {
            int64_t items = 0;
            if (used_[sl]) {
                const MeshArgs &m = solves_[sl].m;
                items = k == kTabFace ? ((int64_t)(m.imax + 1) * m.jmax + (int64_t)m.imax * (m.jmax + 1)) * n
                                      : (int64_t)m.ncell;
            }
            pk[sl + 1] = pk[sl] + (int)((items + quanta[k] - 1) / quanta[k]);
        }
        if (pk[ns] > tab_cap[k]) return ADERDG_ECAPACITY;
    }
    int64_t mx = 0;
    for (int k = 0; k < kNumTabs; k++) mx = std::max<int64_t>(mx, p[k * (ns + 1) + ns]);
    if (staging_.put(prefix_dev_, p, sizeof(int) * kNumTabs * (ns + 1), st_) != hipSuccess) return ADERDG_EHIP;
    if (launch_tables(b, prefix_dev_, ns, quanta, (int)std::max<int64_t>(mx, 1), st_) != hipSuccess) return ADERDG_EHIP;
    return 0;
}

int Group::add(const MeshArgs &m, const SolveState &st0, int *slot)
{
    int sl = -1;
    for (int k = 0; k < caps.slots; k++)
        if (!used_[k]) {
            sl = k;
            break;
        }
    if (sl < 0) return ADERDG_EFULL;
    const int64_t nfx = (int64_t)(m.imax + 1) * m.jmax, nfy = (int64_t)m.imax * (m.jmax + 1);
    const int64_t nlh = lh_slots(pde, m.ncell, caps.lh_fraction);
    const int64_t ac = span(m.ncell), ax = span(nfx), ay = span(nfy);
    int64_t c0 = 0, x0 = 0, y0 = 0, l0 = 0;
    if (!cell_alloc_.alloc(ac, &c0)) return ADERDG_EFULL;
    if (!fx_alloc_.alloc(ax, &x0)) {
        cell_alloc_.release(c0, ac);
        return ADERDG_EFULL;
    }
    if (!fy_alloc_.alloc(ay, &y0)) {
        cell_alloc_.release(c0, ac);
        fx_alloc_.release(x0, ax);
        return ADERDG_EFULL;
    }
    if (!lh_alloc_.alloc(nlh, &l0)) {
        cell_alloc_.release(c0, ac);
        fx_alloc_.release(x0, ax);
        fy_alloc_.release(y0, ay);
        return ADERDG_EFULL;
    }
    SolveArgs &sa = solves_[sl];
    sa = SolveArgs{};
    sa.m = m;
    sa.cell0 = (int)c0;
    sa.fx0 = (int)x0;
    sa.fy0 = (int)y0;
    sa.lh0 = (int)l0;
    sa.lhcap = (int)nlh;
    used_[sl] = 1;
    nslot_ = std::max(nslot_, sl + 1);
    cells_in_use_ += m.ncell;
    SolveArgs *sad = const_cast<SolveArgs *>(b.solves);
    int *cs = const_cast<int *>(b.cell_solve);
    if (staging_.put(sad + sl, &sa, sizeof(SolveArgs), st_) != hipSuccess ||
        staging_.put(b.state + sl, &st0, sizeof(SolveState), st_) != hipSuccess ||
        hipMemsetD32Async((hipDeviceptr_t)(cs + c0), sl, (size_t)m.ncell, st_) != hipSuccess ||
        hipMemsetAsync(b.rc + c0, 0, sizeof(int) * m.ncell, st_) != hipSuccess ||
        hipMemsetAsync(b.rc_final + c0, 0, sizeof(int) * m.ncell, st_) != hipSuccess ||
        hipMemsetAsync(b.status[0] + c0, 0, sizeof(int) * m.ncell, st_) != hipSuccess ||
        hipMemsetAsync(b.status[1] + c0, 0, sizeof(int) * m.ncell, st_) != hipSuccess ||
        hipMemsetAsync(b.lh_count + sl, 0, sizeof(int), st_) != hipSuccess)
        return ADERDG_EHIP;
    int r = rebuild_tables();
    if (r) return r;
    *slot = sl;
    return 0;
}

int Group::remove(int sl)
{
    if (!used(sl)) return ADERDG_EINVAL;
    const MeshArgs &m = solves_[sl].m;
    SolveState off{};  // inactive, no seed pending
    if (staging_.put(b.state + sl, &off, sizeof(SolveState), st_) != hipSuccess) return ADERDG_EHIP;
    cell_alloc_.release(solves_[sl].cell0, span(m.ncell));
    fx_alloc_.release(solves_[sl].fx0, span((int64_t)(m.imax + 1) * m.jmax));
    fy_alloc_.release(solves_[sl].fy0, span((int64_t)m.imax * (m.jmax + 1)));
    lh_alloc_.release(solves_[sl].lh0, solves_[sl].lhcap);
    cells_in_use_ -= m.ncell;
    used_[sl] = 0;
    while (nslot_ > 0 && !used_[nslot_ - 1]) nslot_--;
    return rebuild_tables();
}

int Group::upload_uh(int sl, const double *uh, int parity, hipStream_t cs)
{
    if (!used(sl)) return ADERDG_EINVAL;
    const size_t blk = (size_t)V * n * n;
    const size_t c0 = solves_[sl].cell0, nc = solves_[sl].m.ncell;
    if (hipMemcpyAsync(b.uh[parity] + c0 * blk, uh, nc * blk * sizeof(double), hipMemcpyHostToDevice,
                       cs ? cs : st_) != hipSuccess)
        return ADERDG_EHIP;
    return 0;
}

int Group::download_uh(int sl, double *uh, int parity, hipStream_t cs)
{
    if (!used(sl)) return ADERDG_EINVAL;
    const size_t blk = (size_t)V * n * n;
    const size_t c0 = solves_[sl].cell0, nc = solves_[sl].m.ncell;
    if (hipMemcpyAsync(uh, b.uh[parity] + c0 * blk, nc * blk * sizeof(double), hipMemcpyDeviceToHost,
                       cs ? cs : st_) != hipSuccess)
        return ADERDG_EHIP;
    return 0;
}

int Group::write_state(int sl, const SolveState &st)
{
    if (!used(sl)) return ADERDG_EINVAL;
    if (staging_.put(b.state + sl, &st, sizeof(SolveState), st_) != hipSuccess) return ADERDG_EHIP;
    return 0;
}

int Group::read_states_async(SolveState *pinned_out)
{
    if (nslot_ == 0) return 0;
    if (hipMemcpyAsync(pinned_out, b.state, sizeof(SolveState) * nslot_, hipMemcpyDeviceToHost, st_) != hipSuccess)
        return ADERDG_EHIP;
    return 0;
}

int Group::read_states(SolveState *out)
{
    if (nslot_ == 0) return 0;
    if (hipMemcpyAsync(out, b.state, sizeof(SolveState) * nslot_, hipMemcpyDeviceToHost, st_) != hipSuccess ||
        hipStreamSynchronize(st_) != hipSuccess)
        return ADERDG_EHIP;
    return 0;
}

hipError_t Group::launch(int part) const
{
    return launch_part(pde, N, variant, upd, b, st_, part, grid, caps.slots);
}

hipError_t Group::launch_on(int part, hipStream_t s, bool shared) const
{
    if (shared && part == kPartPredictor && variant == kPredMfma && cus > 0 && grid[kTabPred] > cus) {
        int g[kNumTabs];
        for (int k = 0; k < kNumTabs; k++) g[k] = grid[k];
        g[kTabPred] = cus;
        return launch_part(pde, N, variant, upd, b, s, part, g, caps.slots);
    }
    return launch_part(pde, N, variant, upd, b, s, part, grid, caps.slots);
}

}  // namespace aderdg
