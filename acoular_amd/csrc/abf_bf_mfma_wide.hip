// Instantiations of the wide streamed split-bf16 MFMA beamformer (abf_mfma_wide.h) for NK = Mp / 32
// blocks of 32 microphones: the build compiles this file once per NK in 2 .. 16 with
// -DABF_ST_NK=<NK> (one object per NK, built in parallel), for every steering weight kind
// (0 none, 1 r, 2 1/r).
#include "abf_mfma_wide.h"

#ifndef ABF_ST_NK
#error "compile with -DABF_ST_NK=<2..16>"
#endif

namespace abf {

template __global__ void bf_mfma_wide_kernel<ABF_ST_NK, 0>(const MfJob *, int);
template __global__ void bf_mfma_wide_kernel<ABF_ST_NK, 1>(const MfJob *, int);
template __global__ void bf_mfma_wide_kernel<ABF_ST_NK, 2>(const MfJob *, int);

} // namespace abf
