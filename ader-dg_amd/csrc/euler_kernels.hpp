// euler_kernels.hpp -- launch interface of the element-step kernels (2-D Euler
// and Baer-Nunziato), implemented in step_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "operators.hpp"

namespace aderdg {

constexpr int kBlock = 256;        // spread / dt workgroup size
constexpr int kDtGrid = 256;       // workgroups of the time-step reduction
constexpr int kSubcellGrid = 4096; // persistent workgroups of the subcell recompute

enum { kPdeEuler = 0, kPdeBn = 1 };
enum { kBcDmr = 0, kBcPeriodic = 1, kBcTransmissive = 2 };
enum { kCntList = 0, kCntTroubled = 1, kCntDtBlocks = 2, kNumCounters = 4 };
enum { kPartDt = 0, kPartPredictor = 1, kPartUpdate = 2, kPartSpread = 3, kPartSubcell = 4, kPartAdvance = 5,
       kPartSeed = 6, kPartFace = 7 };

// Mesh and equation constants, passed by value.
struct MeshArgs {
    int pde, imax, jmax, ncell, bc_mode, wrap;
    double dx, dy, gamma, cflpnpm;
    double eos[4];             // Baer-Nunziato: GAMMA_S, GAMMA_G, PI_S, PI_G
    double dmr_a, dmr_b;       // DMR shock position x_sw = dmr_a + dmr_b * t (Euler/riemann_problem.f90:1-8)
    double qa[4], qb[4];       // DMR inflow / pre-shock conserved states
};

// Device buffers of one step (cur = state at the start of the step).
struct StepBufs {
    const Operators *op;
    double *scal;  // [0] dt, [1] time, [2] tend
    const double *uh_cur;
    double *uh_next, *lduh, *fq, *ff;
    double *xflux, *yflux;  // [side][v][p][face]: side 0 D- (Euler: the flux), side 1 D+ (Baer-Nunziato)
    double *lmin_old, *lmax_old, *lmin, *lmax, *cell_dt, *dt_partial;
    int *rc, *rc_final, *list, *counters;
    const int *status_cur;
    int *status_next;
    const double *lh_cur;
    double *lh_next;
};

// Outputs of the standalone predictor (reference layouts per cell; shi only for Baer-Nunziato).
struct TestBufs {
    const double *uh;
    double *qhi, *fhi, *qbnd, *fbnd, *shi;
};

hipError_t launch_part(int N, const MeshArgs &m, const StepBufs &b, hipStream_t st, int part);
hipError_t launch_dt(const MeshArgs &m, const StepBufs &b, hipStream_t st, double time, double tend, int use_host);
hipError_t launch_test_predictor(int N, const MeshArgs &m, const StepBufs &b, const TestBufs &t, hipStream_t st);

}  // namespace aderdg
