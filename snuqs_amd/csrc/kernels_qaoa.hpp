// kernels_qaoa.hpp -- batched QAOA MaxCut (kernels_qaoa.hip).
#pragma once

#include "sq_common.hpp"

namespace sq {

int qaoa_create(void **handle, sq_dtype_t dt, uint32_t ninst, const uint32_t *nq,
                const uint32_t *nedges, const uint32_t *edges, uint32_t p);
int qaoa_run(void *handle, const double *gammas, const double *betas, int use_graph,
             double *out_host, hipStream_t st);
int qaoa_stats(void *handle, uint64_t *launches, double *state_bytes, double *traffic_bytes);
int qaoa_destroy(void *handle);

}  // namespace sq
