// percell.hpp -- per-cell forms of the reference's element routines (percell.hip): the
// volume integral, surface integral, element update and DMP of ncells independent cells
// whose arrays are given in the reference's per-cell layouts, behind the C-ABI entries
// aderdg_volume_integral_cells, aderdg_surface_integral_cells, aderdg_element_update_cells
// and aderdg_dmp_cells.
#pragma once
#include <hip/hip_runtime.h>

#include "operators.hpp"
#include "step_kernels.hpp"

namespace aderdg {

// Device arrays of one per-cell call, each in the reference layout of one cell repeated
// ncells times (cell-major): lduh, luh, shi (V,n,n); fhi (V,2,n,n); fl, fr, gb, gt (V,n);
// lmin, lmax (V); result one int per cell; work V (s^2 + n^2) doubles per cell (DMP).
struct CellBufs {
    double *lduh = nullptr, *luh = nullptr;
    const double *fhi = nullptr, *shi = nullptr;
    const double *fl = nullptr, *fr = nullptr, *gb = nullptr, *gt = nullptr;
    const double *lmin = nullptr, *lmax = nullptr;
    int *result = nullptr;
    double *work = nullptr;
};

enum { kCellVolume = 0, kCellSurface = 1, kCellUpdate = 2, kCellDmp = 3 };

// One per-cell routine (kCell*) over ncells cells; m gives the PDE, mesh spacing and
// equation of state, op the operators (device pointer), dt the element update's time step,
// argmax the DMP's tolerance floor.
hipError_t launch_cells(int kind, const MeshArgs &m, const Operators *op, int N, int ncells, const CellBufs &c,
                        double dt, double argmax, hipStream_t st);

}  // namespace aderdg
