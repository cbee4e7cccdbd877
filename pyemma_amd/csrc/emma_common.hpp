// Shared helpers of the MI355X (gfx950) discretization kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <type_traits>

#include "../../include/emma_amd.h"

namespace emma {

// ------------------------------------------------------------------
// host-side error state
// ------------------------------------------------------------------
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
int check_launch(const char *what);  // EMMA_OK or EMMA_ERR_HIP (records the HIP message)

#define EMMA_REQUIRE(cond, ...)              \
    do {                                     \
        if (!(cond)) {                       \
            ::emma::set_error(__VA_ARGS__);  \
            return EMMA_ERR_INVALID_ARG;     \
        }                                    \
    } while (0)

#define EMMA_TRY(expr)            \
    do {                          \
        int _rc = (expr);         \
        if (_rc != EMMA_OK) return _rc; \
    } while (0)

// ------------------------------------------------------------------
// workspace carving: sub-buffers aligned to 256 bytes
// ------------------------------------------------------------------
struct Carver {
    char *base;
    size_t off;
    explicit Carver(void *p) : base(static_cast<char *>(p)), off(0) {}
    template <typename T>
    T *take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T *p = reinterpret_cast<T *>(base ? base + off : nullptr);
        off += count * sizeof(T);
        return p;
    }
    size_t bytes() const { return (off + 255) & ~size_t(255); }
};

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ------------------------------------------------------------------
// device helpers
// ------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// float -> bf16 bit pattern, round to nearest even (finite inputs).
__device__ __forceinline__ uint32_t bf16_bits_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf16_bits_to_f32(uint32_t h) { return __uint_as_float(h << 16); }

// Split a = hi + lo (+ r, |r| <= 2^-16 |a|) with hi = bf16(a), lo = bf16(a - hi).
// a - hi is exact in FP32.
__device__ __forceinline__ void split_bf16(float a, uint32_t &hi, uint32_t &lo) {
    hi = bf16_bits_rne(a);
    float r = a - bf16_bits_to_f32(hi);
    lo = bf16_bits_rne(r);
}

__device__ __forceinline__ uint32_t pack2(uint32_t lo16, uint32_t hi16) { return lo16 | (hi16 << 16); }

// Two values at once on gfx950's v_cvt_pk_bf16_f32 (round to nearest even):
// hi = packed bf16(a0), bf16(a1); lo = packed bf16(a0 - hi0), bf16(a1 - hi1).
// Same hi/lo as split_bf16 for finite inputs, in 6 instructions per pair.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split_bf16x2(float a0, float a1, uint32_t &hi, uint32_t &lo) {
    const f32x2_t v = {a0, a1};
    hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
    const f32x2_t hf = {__uint_as_float(hi << 16), __uint_as_float(hi & 0xffff0000u)};
    const f32x2_t r = v - hf;
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2_t));
}

// f(std::integral_constant<int, K>()) for K = B .. E-1, in order: loop
// bodies whose register-array indices must be compile-time constants.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>());
        static_for<B + 1, E>(f);
    }
}

// 64-lane butterfly sum (every lane ends with the same value).
__device__ __forceinline__ double wave_butterfly_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

}  // namespace emma
