// kernels_cbatch.hpp -- circuits of uneven qubit counts as one launch
// sequence (kernels_cbatch.hip).
#pragma once

#include "sq_common.hpp"

namespace sq {

int cbatch_create(void **handle, sq_dtype_t dt, uint32_t ncirc, const uint32_t *nq,
                  const uint32_t *ng, const uint32_t *nt, const uint32_t *targets,
                  const double *matrices);
int cbatch_run(void *handle, int use_graph, hipStream_t st);
int cbatch_copy_state(void *handle, uint32_t circuit, void *dst, uint64_t nelems, hipStream_t st);
int cbatch_stats(void *handle, uint64_t *launches, double *state_bytes, double *traffic_bytes);
int cbatch_destroy(void *handle);

}  // namespace sq
