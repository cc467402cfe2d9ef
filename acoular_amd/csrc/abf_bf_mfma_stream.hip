// Instantiations of the streamed split-bf16 MFMA beamformer (abf_mfma_stream.h) for NK = Mp / 32
// blocks of 32 microphones: the build compiles this file once per NK in 2 .. 8 with
// -DABF_ST_NK=<NK> (one object per NK, built in parallel),
// for every steering weight kind (0 none, 1 r, 2 1/r), one item per workgroup and persistent.
#include "abf_mfma_stream.h"

#ifndef ABF_ST_NK
#error "compile with -DABF_ST_NK=<2..8>"
#endif

namespace abf {

template __global__ void bf_mfma_stream_kernel<ABF_ST_NK, 0, false>(const MfJob *, int, int64_t);
template __global__ void bf_mfma_stream_kernel<ABF_ST_NK, 1, false>(const MfJob *, int, int64_t);
template __global__ void bf_mfma_stream_kernel<ABF_ST_NK, 2, false>(const MfJob *, int, int64_t);
template __global__ void bf_mfma_stream_kernel<ABF_ST_NK, 0, true>(const MfJob *, int, int64_t);
template __global__ void bf_mfma_stream_kernel<ABF_ST_NK, 1, true>(const MfJob *, int, int64_t);
template __global__ void bf_mfma_stream_kernel<ABF_ST_NK, 2, true>(const MfJob *, int, int64_t);

} // namespace abf
