// group.hpp -- host side of a solve group: the device arena of one (PDE, N),
// the solves placed in it, and the work tables the kernels loop over.
//
// A group owns flat device arrays sized by its capacities (cells, x-faces,
// y-faces, solve slots).  Adding a solve takes first-fit ranges of cells and
// faces and a slot, writes the solve's SolveArgs / SolveState and rebuilds the
// work tables on the device; removing it returns the ranges.  Neither changes
// a kernel's launch configuration, so a captured step graph stays valid.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <vector>

#include "operators.hpp"
#include "step_kernels.hpp"

namespace aderdg {

// Pinned host ring for small host-to-device uploads on one stream: a chunk is
// reused only after the event recorded behind its copy has completed, so an
// upload never waits for the kernels queued before it.
class Staging {
public:
    int init(size_t bytes);
    void release();
    // Copy n bytes of src to device dst on stream st through the ring.
    hipError_t put(void *dst, const void *src, size_t n, hipStream_t st);

private:
    struct Pending {
        size_t start, end;
        hipEvent_t ev;
    };
    char *buf_ = nullptr;
    size_t cap_ = 0, pos_ = 0;
    std::vector<Pending> pending_;  // oldest first
    std::vector<hipEvent_t> free_;
};

// Solves' cell and face ranges start on multiples of kAlign items and take multiples of it:
// 16 doubles, one 128-byte line of every [v][cell] / [v][p][face] plane, so that a workgroup of
// 16 cells (the matrix-core predictor's) writes whole lines of the face-extrapolation planes.
constexpr int64_t kAlign = 16;
inline int64_t align_up(int64_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator of index ranges in [0, cap), coalescing on release.
class RangeAlloc {
public:
    void reset(int64_t cap);
    bool alloc(int64_t len, int64_t *start);
    void release(int64_t start, int64_t len);
    int64_t largest_free() const;
    int64_t capacity() const { return cap_; }

private:
    int64_t cap_ = 0;
    std::map<int64_t, int64_t> free_;  // start -> length
};

struct GroupCaps {
    int64_t cells, xfaces, yfaces;
    int slots;
    // limited-cell pool: a solve of C cells gets ceil(lh_fraction * C) slots of subcell
    // averages (1: one per cell, a solve can never run out)
    double lh_fraction = 1.0;
};

class Group {
public:
    int pde = 0, N = 0, V = 0, n = 0, s = 0;
    GroupCaps caps{};
    Operators ops{};
    int variant = kPredLine;  // predictor kernel variant (step_kernels.hpp)
    int upd = kUpdShared;     // update kernel variant
    int quanta[kNumTabs] = {};
    int grid[kNumTabs] = {};
    int cus = 0;  // compute units of the device (0: unknown)
    int64_t tab_cap[kNumTabs] = {};
    StepBufs b{};

    ~Group() { release(); }
    // Allocate the arena.  persist_blocks: workgroups of the table-driven kernels
    // (0: one per table entry of a full arena); variant: predictor kernel variant,
    // or -1 for the autotuner's choice (tune.hip).
    int init(int pde, int N, const GroupCaps &caps, const Operators &ops, int persist_blocks, hipStream_t st,
             int variant = -1, int upd = -1);
    void release();

    // Place a solve; returns 0, or ADERDG_EFULL when no slot or range is free.
    int add(const MeshArgs &m, const SolveState &st0, int *slot);
    int remove(int slot);
    const MeshArgs &mesh(int slot) const { return solves_[slot].m; }
    int cell0(int slot) const { return solves_[slot].cell0; }
    const SolveArgs &args(int slot) const { return solves_[slot]; }
    // items a solve's range takes: aligned in arenas of several slots
    int64_t span(int64_t n) const { return caps.slots > 1 ? align_up(n) : n; }
    bool used(int slot) const { return slot >= 0 && slot < caps.slots && used_[slot]; }
    int nslot() const { return nslot_; }
    int64_t cells_in_use() const { return cells_in_use_; }

    // Host <-> device for one solve; uh in the reference layout uh(V,n,n,IMAX,JMAX).
    // Solution copies on stream cs (default: the group's stream).
    int upload_uh(int slot, const double *uh, int parity, hipStream_t cs = nullptr);
    int download_uh(int slot, double *uh, int parity, hipStream_t cs = nullptr);
    int write_state(int slot, const SolveState &st);
    int read_states(SolveState *out);                  // all nslot() states (synchronous)
    int read_states_async(SolveState *pinned_out);     // enqueued on the group's stream
    hipStream_t stream() const { return st_; }

    hipError_t launch(int part) const;
    // The same launch on another stream.  shared: the launch is one branch of a graph that
    // steps several arenas side by side; the matrix-core predictor then runs one workgroup
    // per CU (one wave per SIMD) instead of filling the CUs, so that the other arenas'
    // passes run beside it on every CU (at one wave per SIMD it keeps about 80% of its rate:
    // the Baer-Nunziato ensemble step 47.6 -> 46.4 ms)
    hipError_t launch_on(int part, hipStream_t s, bool shared = false) const;
    int dt_grid() const { return grid[kTabCell] < kDtGrid ? grid[kTabCell] : kDtGrid; }
    // Bytes of device memory per cell / x-face / y-face / solve slot / limited-cell pool
    // slot of a (PDE, N).
    static void footprint(int pde, int N, int64_t *per_cell, int64_t *per_xface, int64_t *per_yface,
                          int64_t *per_slot, int64_t *per_lh);
    // Pool slots of a solve of ncell cells, and of a group, at a pool fraction.
    static int64_t lh_slots(int pde, int64_t ncell, double fraction);
    static int64_t lh_pool(int pde, int64_t cells, int slots, double fraction);

private:
    hipStream_t st_ = nullptr;
    std::vector<void *> allocs_;
    std::vector<SolveArgs> solves_;
    std::vector<char> used_;
    RangeAlloc cell_alloc_, fx_alloc_, fy_alloc_, lh_alloc_;
    int nslot_ = 0;
    int64_t cells_in_use_ = 0;
    int *prefix_dev_ = nullptr;
    std::vector<int> prefix_host_;
    Staging staging_;
    int rebuild_tables();
    template <class T>
    int dalloc(T **p, size_t count);
};

// The autotuner's predictor and update variants for a (PDE, N) and launch mode on
// the current device (tune.hip); times the variants the first time it is asked.
void tune_kernels(int pde, int N, int persist_blocks, const Operators &ops, hipStream_t st, int *pred, int *upd);

}  // namespace aderdg
