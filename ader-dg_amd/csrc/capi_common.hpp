// capi_common.hpp -- helpers shared by the C-ABI translation units: error
// reporting, descriptor validation and the per-solve constants.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/aderdg.h"
#include "operators.hpp"
#include "step_kernels.hpp"

namespace aderdg {

// Record the message returned by aderdg_last_error() and return code.
int fail(int code, const char *fmt, ...);

#define HIPCHK(expr)                                                                                           \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess)                                                                                  \
            return ::aderdg::fail(ADERDG_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

// Validate a solve descriptor (who names the caller in messages); 0 or ADERDG_EINVAL.
int check_desc(const aderdg_desc *d, const char *who);
// Mesh, boundary and equation constants of a validated descriptor.
MeshArgs mesh_from_desc(const aderdg_desc *d);
// Operators of degree d->N, with the host matrices of d->ops where given.
int ops_from_desc(const aderdg_desc *d, Operators *o, const char *who);

}  // namespace aderdg
