// 16-bit storage conversions shared by the mixed-precision kernels:
// round-to-nearest-even FP32 -> BF16 / FP16 and exact widening back, equal
// bit for bit to the oracle's integer restatement (or_round16 / or_widen16
// in oracle/stencil_oracle.c) on finite values.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace st {

// BF16 rounding: the gfx950 conversion (v_cvt_pk_bf16_f32) rounds to nearest
// even and keeps FP32 subnormals, which is the integer restatement
// (x + 0x7fff + lsb) >> 16 for every non-NaN input, infinities and overflow
// included; a NaN stays a NaN (the restatement sets its quiet bit).
__device__ __forceinline__ uint16_t to_bf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// FP16: the hardware conversions.  v_cvt_f16_f32 rounds to nearest even
// with FP16 subnormals kept (the default float mode of HIP kernels), which is
// the integer restatement's rounding for every finite input and for
// infinities; v_cvt_f32_f16 widens exactly.  (NaN payloads may differ from
// the restatement, which returns a canonical quiet NaN.)
__device__ __forceinline__ uint16_t to_f16(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }

__device__ __forceinline__ float from_f16(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

template <bool F16> __device__ __forceinline__ uint16_t rnd16(float f) { return F16 ? to_f16(f) : to_bf16(f); }
template <bool F16> __device__ __forceinline__ float wid16(uint16_t h) { return F16 ? from_f16(h) : from_bf16(h); }

}  // namespace st
