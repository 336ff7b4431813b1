// mlp.hpp -- internal interface between the network (mlp.hip) and the solver context (solver.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "mstcfd.h"

namespace mst {
int fail(int code, const std::string &msg);
// de-normalised outputs dY [rows x n_out] of device features dX [rows x n_in]; the caller's stream
// is drained before and the network's stream after
int mlp_forward_device(mst_mlp *q, const float *dX, int64_t rows, float *dY, hipStream_t stream);
int mlp_dims(const mst_mlp *q, int *n_in, int *n_out);
}  // namespace mst
