// step_kernels.hpp -- launch interface of the element-step kernels (2-D Euler
// and Baer-Nunziato), implemented in step_kernels.hip.
//
// The kernels work on a group: any number of independent solves of one
// (PDE, N), concatenated in flat device arrays.  A solve is described by
// SolveArgs (static: mesh, boundaries, equation of state, offsets into the
// flat arrays) and SolveState (dynamic: dt, time, ping-pong parity, active
// flag).  Workgroups take their work from device-resident tables of
// (solve, first item) entries that never straddle two solves, so everything a
// workgroup reads about its solve is uniform.  The single-instance solver is a
// group of one solve.
#pragma once
#include <hip/hip_runtime.h>

#include "ic_spec.hpp"
#include "operators.hpp"

namespace aderdg {

constexpr int kBlock = 256;        // spread / dt / face workgroup size
constexpr int kDtGrid = 256;       // workgroups of the time-step reduction
constexpr int kSubcellGrid = 4096; // persistent workgroups of the subcell recompute

enum { kPdeEuler = 0, kPdeBn = 1, kPdeAd = 2 };
enum { kBcDmr = 0, kBcPeriodic = 1, kBcTransmissive = 2, kBcSides = 3 };
// boundary code of one side (the reference's params.ini convention, Euler/params.ini:8-11)
enum { kSideInflow = 0, kSideTransmit = 1, kSideReflect = 2, kSideWrap = 3 };
enum { kCntList = 0, kCntTroubled = 1, kCntDtBlocks = 2, kNumCounters = 4 };
// kPartRanges (the DMP bounds of the step) runs before every kPartPredictor
enum { kPartDt = 0, kPartPredictor = 1, kPartUpdate = 2, kPartSpread = 3, kPartSubcell = 4, kPartAdvance = 5,
       kPartSeed = 6, kPartFace = 7, kPartFlip = 8, kPartSeedDone = 9, kPartRanges = 10 };
// work tables: items per entry are cells (predictor, update, cell) or face points (face)
enum { kTabPred = 0, kTabUpd = 1, kTabFace = 2, kTabCell = 3, kNumTabs = 4 };
// device meta words: [0, kNumTabs) table lengths, then the number of solve slots in use
enum { kMetaSolves = kNumTabs, kNumMeta = kNumTabs + 1 };

// Mesh and equation constants of one solve.
struct MeshArgs {
    int pde, imax, jmax, ncell, bc_mode;
    int wrap;                  // neighbour table: bit 0 wraps in x, bit 1 in y (periodic sides)
    int ad_fixed;              // advection-diffusion: predictor flux (0 viscous only, 1 fused)
    int ad_bc[4];              // advection-diffusion bC(1..4): 1 Neumann, 2 Dirichlet, 3 periodic
    int side[4];               // Euler / Baer-Nunziato: left, right, bottom, top code (kSide*); DMR: unused
    int pad;
    double dx, dy, gamma, cflpnpm;
    double adv[2], diff;       // advection-diffusion EQN%cx, EQN%cy, EQN%a
    double xl[2];              // lower-left corner of the domain
    double eos[4];             // Baer-Nunziato: GAMMA_S, GAMMA_G, PI_S, PI_G
    double dmr_a, dmr_b;       // DMR shock position x_sw = dmr_a + dmr_b * t (Euler/riemann_problem.f90:1-8)
    double qa[4], qb[4];       // DMR inflow / pre-shock conserved states
    double qin[9];             // conserved ghost state of kSideInflow sides
};

// Static description of one solve of a group.
struct SolveArgs {
    MeshArgs m;
    int cell0;      // first cell in the group's cell arrays
    int fx0, fy0;   // first x-face ((IMAX+1)*JMAX of them) and y-face (IMAX*(JMAX+1))
    int lh0, lhcap; // the solve's segment of the limited-cell pool: first slot, slots
    int pad;
};

// Dynamic per-solve scalars, written by the kernels.
struct SolveState {
    double dt, time, tend;
    unsigned long long dtmin;  // ordered bits of the smallest time-step candidate, ~0 = none
    int parity;                // buffer (0/1) that holds the current solution
    int active;                // 1 while the solve is stepping
    int seed;                  // 1 until the limiter seeds of a fresh solution are computed
    int steps, max_steps;      // steps taken, and the count that ends the solve
    int stop_at_tend;          // 1: the solve also ends once time >= tend
    int failed;                // 1: a step limited more cells than the solve's pool segment holds
    int pad;
};

// Device buffers of a group.  Cell arrays are strided by G (cell capacity), face
// arrays by GX / GY (x- / y-face capacity).
struct StepBufs {
    const Operators *op;
    const SolveArgs *solves;
    SolveState *state;
    const int2 *tab[kNumTabs];  // (solve, first item) per workgroup-sized piece of work
    const int *meta;
    int G, GX, GY;
    const int *cell_solve;      // solve slot of every cell (compacted-list kernels)
    double *uh[2];              // [cell][j][i][v], the reference's uh(V,n,n,IMAX,JMAX) per cell
    double *lduh;
    double *fq, *ff;            // [face][v][p][cell] face extrapolations of q and F*n
    double *xflux, *yflux;      // [side][v][p][face]: side 0 D- (Euler: the flux), side 1 D+ (Baer-Nunziato)
    double *lmin_old, *lmax_old, *lmin, *lmax, *cell_dt;  // [v][cell], [cell]
    int *rc, *rc_final, *list, *counters;
    int *status[2];
    // The limiter's subcell averages (the reference's Lh(V,s,s) of tLimiter, allocated per
    // limited cell) in a pool of slots: lh[par] + slot * V * s * s, slot = lh_slot[par][cell]
    // for a cell with status[par][cell] = 1.  A solve owns the slots [lh0, lh0 + lhcap); a
    // step takes them in the order its limited cells finish (lh_count[solve] counts them).
    double *lh[2];              // [slot][q][p][v]
    int *lh_slot[2];            // [cell]; -1: no slot was left (the solve has failed)
    int *lh_count;              // [solve slot] pool slots taken since the last step ended
};

// Outputs of the standalone predictor (reference layouts per cell; shi only for Baer-Nunziato).
struct TestBufs {
    const double *uh;
    double *qhi, *fhi, *qbnd, *fbnd, *shi;
};

// Predictor kernels: time-line threads (one thread per spatial node holds the node's
// whole time line; kPredLean: the same mapping with a smaller register footprint) or
// space-time-node threads (one thread per node of the space-time cell) in workgroups
// of 64, 128 or 256.  All give the same bits; the autotuner
// (tune.hip) picks one per (PDE, N).  kPredMfma (N <= 3, product build only): time-line
// threads, one cell per 16 lanes, the contractions, volume integral and boundary
// extrapolation on the matrix cores; within rounding of the others.
// Advection-diffusion has the time-line form only.
enum {
    kPredLine = 0, kPredNode64 = 1, kPredNode128 = 2, kPredNode256 = 3, kPredLean = 4, kPredMfma = 5,
    kNumPredVariants = 6
};
bool pred_variant_valid(int pde, int N, int variant);
// Update kernels: kUpdRecompute (k_update: each thread forms the first passes of the
// subcell and Lobatto maps it needs; LDS atomics reduce the ranges) or kUpdShared
// (k_update_shared: the first passes shared in LDS, partials reduced by the cell's
// threads).  The same results; the autotuner picks one per (PDE, N).
enum { kUpdRecompute = 0, kUpdShared = 1, kNumUpdVariants = 2 };
int pred_cells_per_block(int pde, int N, int variant);

// Workgroups of the kernel that works through table `tab` that fit on one CU at once
// (0: not known, or not a table kernel of this PDE).  A persistent grid larger than
// this times the CU count leaves its last workgroups to a second round on the CUs.
int resident_blocks_per_cu(int pde, int N, int variant, int upd, int tab);

// Items (cells or face points) per work-table entry of each table, for a (PDE, N)
// and predictor variant.
void table_quanta(int pde, int N, int variant, int quanta[kNumTabs]);
// the advection-diffusion kernels (ad_kernels.hip)
void table_quanta_ad(int N, int quanta[kNumTabs]);
hipError_t launch_part_ad(int N, const StepBufs &b, hipStream_t st, int part, const int grid[kNumTabs]);

// One part of the step for every solve of the group.  grid: workgroups per table-driven
// kernel (they loop over the table); nslot_cap: solve slots of the group.
hipError_t launch_part(int pde, int N, int variant, int upd, const StepBufs &b, hipStream_t st, int part,
                       const int grid[kNumTabs], int nslot_cap);
// calc_time_step of every active solve; use_host: time and tend given here instead of the solve state.
// grid: workgroups (at most kDtGrid; the cell table's length is enough, and fewer
// workgroups means fewer finalisation tickets on one counter).
hipError_t launch_dt(const StepBufs &b, hipStream_t st, double time, double tend, int use_host, int grid = kDtGrid);
// Rebuild the work tables from per-table prefix counts of entries over the solve slots
// (prefix[k * (nslot + 1) + s], device memory); max_entries bounds every table's length.
hipError_t launch_tables(const StepBufs &b, const int *prefix, int nslot, const int quanta[kNumTabs], int max_entries,
                         hipStream_t st);
// Troubled-cell seeding of solve `slot` from its current solution and the formula it
// was projected from (the end of initialize): status and Lh of the current parity.
// Seeded cells take pool slots as a step's limited cells do (kPartSeedDone ends the count).
hipError_t launch_seed_ic(int pde, int N, const StepBufs &b, const IcSpec &ic, int slot, int ncell, hipStream_t st);
// Initial state of solve `slot` into uh[parity] from a formula (ic_spec.hpp).
hipError_t launch_init(int pde, int N, const StepBufs &b, const IcSpec &ic, int slot, int ncell, int parity,
                       hipStream_t st);
hipError_t launch_test_predictor(int N, int variant, const MeshArgs &m, const StepBufs &b, const TestBufs &t,
                                 double dt, hipStream_t st);

}  // namespace aderdg
