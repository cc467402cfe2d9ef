// Instantiations of the split-bf16 (3-term) MFMA beamformer kernels: CW column tiles per U fragment,
// TT row tiles per unit, software-pipelined or not, steering-generation variants
// (abf_mfma_impl.h).
#include "abf_mfma_impl.h"

namespace abf {

template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 0, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 0, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 4, 0, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 0, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 0, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 0, 1, true>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 0, 1, true>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 0, 2, true>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<kPrBF16X3, 0>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 1, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 1, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 4, 1, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 1, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 1, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 1, 1, true>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 1, 1, true>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 1, 2, true>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<kPrBF16X3, 1>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 2, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 2, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 4, 2, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 2, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 2, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 2, 1, true>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 2, 1, true>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 2, 2, true>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<kPrBF16X3, 2>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 3, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 3, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 4, 3, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 3, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 3, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 3, 1, true>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 3, 1, true>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 3, 2, true>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<kPrBF16X3, 3>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 4, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 4, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 4, 4, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 4, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 4, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 4, 1, true>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 4, 1, true>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 4, 2, true>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<kPrBF16X3, 4>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 5, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 5, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 4, 5, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 5, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 5, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 5, 1, true>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 2, 5, 1, true>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrBF16X3, 1, 5, 2, true>(const MfJob *, int);
template __global__ void bf_mfma_ws_kernel<kPrBF16X3, 5>(const MfJob *, int);

} // namespace abf
