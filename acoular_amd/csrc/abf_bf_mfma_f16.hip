// Instantiations of the f16x2 MFMA beamformer kernels (U rounded once to fp16 with a power-of-two
// scale per frequency, the steering matrix split hi+lo in fp16): CW column tiles per U fragment,
// TT row tiles per unit, steering-generation variants (abf_mfma_impl.h).  Only the alternating
// kernels exist in this form.
#include "abf_mfma_impl.h"

namespace abf {

template __global__ void bf_mfma_kernel<kPrF16X2, 1, 0, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 2, 0, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 4, 0, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 1, 0, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 2, 0, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 1, 1, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 2, 1, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 4, 1, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 1, 1, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 2, 1, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 1, 2, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 2, 2, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 4, 2, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 1, 2, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 2, 2, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 1, 3, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 2, 3, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 4, 3, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 1, 3, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 2, 3, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 1, 4, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 2, 4, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 4, 4, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 1, 4, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 2, 4, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 1, 5, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 2, 5, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 4, 5, 1, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 1, 5, 2, false>(const MfJob *, int);
template __global__ void bf_mfma_kernel<kPrF16X2, 2, 5, 2, false>(const MfJob *, int);

} // namespace abf
