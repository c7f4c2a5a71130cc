// batch.hpp -- the batched engine behind aderdg_batch_* (batch.hip), also
// driven directly by the serving loop (server.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <vector>

#include "../../include/aderdg.h"
#include "group.hpp"

namespace aderdg {

class Batch {
public:
    struct Handle {
        int group, slot;
    };
    int device = 0, persist = 0;
    hipStream_t stream = nullptr;  // kernels, state traffic
    hipStream_t copy = nullptr;    // solution uploads
    std::vector<std::unique_ptr<Group>> groups;
    std::vector<SolveState *> pinned_states;  // per group, caps.slots entries
    std::map<int, Handle> solves;
    int next_id = 0;
    int64_t reserved = 0;
    hipGraphExec_t graph = nullptr;
    bool graph_stale = true;
    // the step graph forks one branch per arena: arena 0 on `stream`, arena k on branch[k-1]
    // (default priority, like every stream here), joined back into `stream` at the end
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr;
    hipEvent_t upload_done = nullptr, removed = nullptr, states_ready = nullptr;
    // per-kernel timing (steps launched part by part with events instead of the graph)
    struct Timed {
        int group, slot;
        hipEvent_t a, b;
    };
    bool timing = false;
    std::vector<Timed> pending;
    std::vector<hipEvent_t> free_events;
    std::vector<std::vector<double>> ms;        // [group][timing slot]
    std::vector<std::vector<int64_t>> launches;
    int collect_timing();

    int init(int device, int persist);
    void release();
    // lh_fraction: limited-cell pool slots per cell of a solve (GroupCaps)
    int reserve(int pde, int N, int64_t max_cells, int max_solves, double lh_fraction = 1.0);
    // initial state from the host array uh, or computed on the device from ic
    int add(const aderdg_desc *d, const double *uh, const aderdg_ic *ic, double time, double tend, int max_steps,
            int stop_at_tend, int *id);
    int step(int nsteps);
    // Solve states into pinned memory: asynchronous (ready when snapshot_ready()) or synchronous.
    int snapshot_async();
    bool snapshot_ready() const;
    int snapshot();
    const SolveState *state_of(int id) const;  // from the last snapshot
    int download(int id, double *uh);
    // Limiter state of a solve in host layout (see aderdg_get_limiter); outputs may be null.
    int limiter(int id, int32_t *status, int32_t *recompute, double *Lh = nullptr, double *lmin = nullptr,
                double *lmax = nullptr);
    int error(int id, int *err);  // ADERDG_ECAPACITY: the solve ran out of limited-cell pool slots
    // Non-finite DOFs of a solve's current solution and its troubled cells of the last step
    int health(int id, int64_t *nonfinite, int64_t *troubled);
    unsigned long long *health_buf = nullptr;
    int remove(int id);
    int group_of(int pde, int N) const;
    static int64_t footprint(int pde, int N, int imax, int jmax, double lh_fraction = 1.0);
    double lh_fraction(int pde, int N) const;  // of the (pde, N) arena, 1 if none

private:
    int build_graph();
};

}  // namespace aderdg
