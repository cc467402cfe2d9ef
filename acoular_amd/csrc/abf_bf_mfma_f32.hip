// Instantiations of the fp32 MFMA beamformer kernels: CW column tiles per U fragment,
// TT row tiles per unit, software-pipelined or not, steering-generation variants
// (abf_mfma_impl.h).
#include "abf_mfma_impl.h"

namespace abf {

template __global__ void bf_mfma_kernel<kPrF32, 1, 0, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 2, 0, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 4, 0, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 1, 0, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 2, 0, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<kPrF32, 0>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 1, 1, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 2, 1, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 4, 1, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 1, 1, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 2, 1, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<kPrF32, 1>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 1, 2, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 2, 2, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 4, 2, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 1, 2, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 2, 2, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<kPrF32, 2>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 1, 3, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 2, 3, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 4, 3, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 1, 3, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 2, 3, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<kPrF32, 3>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 1, 4, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 2, 4, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 4, 4, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 1, 4, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 2, 4, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<kPrF32, 4>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 1, 5, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 2, 5, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 4, 5, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 1, 5, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF32, 2, 5, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<kPrF32, 5>(const MfJob *, int);

} // namespace abf
