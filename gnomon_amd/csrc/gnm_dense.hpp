// Launchers of the device-side plan builders and the adjoint kernel
// (gnm_dense.hip).  Not part of the C ABI.
#pragma once

#include "gnm_internal.hpp"

namespace gnm {

// int8: limb digits all in [-128, 127] (max |Q| = 127 * 256^(L-1)); else
// BF16 limbs (max |Q| = 2^(8L) - 2^(8(L-1))).
hipError_t launch_block_exponent(const float *src, uint64_t rows, uint64_t ld, uint64_t cols,
                                 uint64_t block, uint32_t L, bool int8, int *out, hipStream_t stream);
// split-BF16 B fragments of nst 64-row stages from dense W [V][K]
hipError_t launch_forward_frag(const float *W, uint64_t V, uint64_t K, uint32_t L, const int *exps,
                               uint32_t wtiles, uint32_t tiles, uint64_t nst, uint16_t *frag,
                               double *scales, hipStream_t stream);
hipError_t launch_dense_program(const uint8_t *flip, uint64_t V, uint64_t Rpad, uint32_t *prog_var,
                                uint32_t *flipmask, uint32_t *prog_cls, hipStream_t stream);
hipError_t launch_pad_weights(const float *W, uint64_t V, uint64_t K, uint64_t Rpad, uint64_t k32,
                              float *out, hipStream_t stream);
hipError_t launch_take_col0(const uint32_t *counts, uint64_t N, uint64_t K, uint32_t *out,
                            hipStream_t stream);
hipError_t launch_adjoint_frag(const float *R, uint64_t N, uint64_t K, uint64_t Kp, uint32_t L,
                               const int *exps, uint32_t tiles, uint64_t nks, uint16_t *frag,
                               double *scales, hipStream_t stream);
hipError_t launch_adjoint(const uint8_t *rows, uint64_t stride, uint64_t V, uint64_t N,
                          const uint8_t *flip, const uint32_t *frag, const double *scales,
                          uint32_t tiles, uint64_t K, uint64_t Kp, uint32_t L, float *dW,
                          hipStream_t stream);

}  // namespace gnm
