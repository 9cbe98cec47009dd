// device.h -- wave-level helpers shared by the gfx950 kernels (64-lane
// wavefronts, DPP row operations, scalar-register copies).
#pragma once
#include <hip/hip_runtime.h>

#include "layout.h"

namespace culda {

namespace {

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Loads and stores through a pointer in the global address space: a plain
// pointer read from a descriptor is generic, and a flat access also counts
// against the LDS counter, so a wait for an LDS read would wait for it too.
#if defined(__HIP_DEVICE_COMPILE__)
#define CULDA_AS1 __attribute__((address_space(1)))
#else
#define CULDA_AS1
#endif
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t load_global(const uint32_t *p) { return *(const CULDA_AS1 uint32_t *)p; }
__device__ __forceinline__ u32x2 load_global2(const uint32_t *p) { return *(const CULDA_AS1 u32x2 *)p; }
__device__ __forceinline__ u32x4 load_global4(const void *p) { return *(const CULDA_AS1 u32x4 *)p; }
__device__ __forceinline__ uint32_t load_global_u16(const uint16_t *p) { return *(const CULDA_AS1 uint16_t *)p; }
__device__ __forceinline__ void store_global_u16(uint16_t *p, uint16_t v) { *(CULDA_AS1 uint16_t *)p = v; }
// Global-address-space view of a generic pointer, and a relaxed
// device-scope atomic add through it (a global_atomic, not a flat one).
template <class T>
__device__ __forceinline__ CULDA_AS1 T *glb(T *p) { return (CULDA_AS1 T *)p; }
__device__ __forceinline__ void atomic_add_global(int32_t *p, int32_t v) {
    __hip_atomic_fetch_add(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The same at base + a 32-bit byte offset (base wave-uniform): the address
// is the scalar base plus a 32-bit vector offset, no 64-bit vector math.
__device__ __forceinline__ u32x2 load_global2_at(const void *base, uint32_t off) {
    return *(const CULDA_AS1 u32x2 *)((const char *)base + off);
}
// ... plus a constant byte offset, added after the 32-bit offset (so it
// can ride in the instruction's immediate field)
template <int IMM>
__device__ __forceinline__ u32x2 load_global2_at_imm(const void *base, uint32_t off) {
    return *(const CULDA_AS1 u32x2 *)((const char *)base + off + IMM);
}
__device__ __forceinline__ uint32_t load_global_at(const void *base, uint32_t off) {
    return *(const CULDA_AS1 uint32_t *)((const char *)base + off);
}
__device__ __forceinline__ uint32_t load_global_u16_at(const void *base, uint32_t off) {
    return *(const CULDA_AS1 uint16_t *)((const char *)base + off);
}
__device__ __forceinline__ void store_global_u16_at(void *base, uint32_t off, uint16_t v) {
    *(CULDA_AS1 uint16_t *)((char *)base + off) = v;
}

typedef double f64x2 __attribute__((ext_vector_type(2)));

// fl(x / d) for FP32 x, d > 0 from r = RN64(1 / d): RN32(RN64(x * r)).  x / d
// is never a midpoint between two FP32 numbers and lies at least 2^-49
// (relative) from every midpoint (x - d m is a nonzero multiple of 2^-47
// for x, d in [1, 2) and a 25-bit m), while RN64(x * r) is within 2^-52
// of x / d, so both round to the same FP32 number -- the correctly rounded
// quotient.  One FP64 multiply instead of an FP32 division sequence.
__device__ __forceinline__ float phi_of(float x, double r) { return (float)((double)x * r); }

// DPP move with zero for lanes whose source is outside the 16-lane row.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xf, 0xf, false));
}

// Copy of a wave-uniform struct held in scalar registers (each dword
// through readfirstlane), so descriptors read inside a loop that also
// stores do not occupy vector registers.
template <class T>
__device__ __forceinline__ T uniform_copy(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "dword-sized struct expected");
    T v = *p;
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 4; ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
    __builtin_memcpy(&v, w, sizeof(T));
    return v;
}

// Lane j's 64-bit value (j wave-uniform), through two scalar reads.
__device__ __forceinline__ int64_t readlane64(int64_t x, int j) {
    const uint64_t u = (uint64_t)x;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), j);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// 64-lane inclusive scans with DPP (row shifts inside each 16-lane row,
// then row_bcast:15 / row_bcast:31 across rows): a few cycles per step
// instead of an LDS round trip per shuffle.  `id` fills lanes without a
// source.  wave_*_all returns the total to every lane.
template <int CTRL, int RM, bool BC>
__device__ __forceinline__ uint32_t dppu(uint32_t old, uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)x, CTRL, RM, 0xf, BC);
}

template <class Op>
__device__ __forceinline__ uint32_t wave_scan_u32(uint32_t x, uint32_t id, Op op) {
    x = op(x, dppu<0x111, 0xf, false>(id, x));
    x = op(x, dppu<0x112, 0xf, false>(id, x));
    x = op(x, dppu<0x114, 0xf, false>(id, x));
    x = op(x, dppu<0x118, 0xf, false>(id, x));
    x = op(x, dppu<0x142, 0xa, false>(id, x));
    x = op(x, dppu<0x143, 0xc, false>(id, x));
    return x;
}

__device__ __forceinline__ uint32_t wave_add_scan_u32(uint32_t x) {
    return wave_scan_u32(x, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}

// ---- 8-lane rows: each 16-lane DPP row holds two of them (lanes 0-7 and
// 8-15 of the DPP row); m = lane & 7 is the lane's place in its row.

// Minimum over each 8-lane row, in every lane of the row: quad_perm
// [1,0,3,2], quad_perm [2,3,0,1], then row_half_mirror (lane m meets lane
// 7 - m, joining the row's two quads).
__device__ __forceinline__ uint32_t row8_min(uint32_t x) {
    x = min(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xf, 0xf, false));
    x = min(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xf, 0xf, false));
    return min(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xf, 0xf, false));
}

// Value of the lane S places below in the same 8-lane row, 0 for m < S:
// row_shr:S over the 16-lane DPP row, after zeroing the lanes m >= 8 - S,
// which are read only by the first S lanes of the upper row.  S = 4 masks
// banks 0 and 2 instead (lanes m < 4 keep the 0 passed as old value).
template <int S>
__device__ __forceinline__ float row8_shr(float x, int m) {
    if (S == 4)
        return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x114, 0xf, 0xA, false));
    return dpp_f<0x110 + S>(m >= 8 - S ? 0.0f : x);
}

// In-block prefixes of 16 products held two per lane by an 8-lane row
// (lane m: product 2m as its factors a * b, product 2m + 1 as the rounded
// value hi): pair sums q_m = fma(a, b, hi) (one rounding), their inclusive
// Kogge-Stone scan P_m (offsets 1, 2, 4), then C_2m+1 = P_m and C_2m =
// fma(a, b, P_m-1) with P_-1 = 0 (so C_0 = fl(a * b)).  C_15 = P_7 is the
// balanced tree of the 8 pair sums, the block total.
__device__ __forceinline__ void prefix16x2(float a, float b, float &lo, float &hi, int m) {
    float P = __builtin_fmaf(a, b, hi);
    P = P + row8_shr<1>(P, m);
    P = P + row8_shr<2>(P, m);
    P = P + row8_shr<4>(P, m);
    lo = __builtin_fmaf(a, b, row8_shr<1>(P, m));
    hi = P;
}

// ---- 4-lane rows (four per 16-lane DPP row); q = lane & 3.

// Minimum over each 4-lane row, in every lane of it.
__device__ __forceinline__ uint32_t quad_min(uint32_t x) {
    x = min(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xf, 0xf, false));
    return min(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xf, 0xf, false));
}

// Value of the lane S places below in the same 4-lane row, 0 for q < S
// (sources that would cross into the next row zeroed first).
template <int S>
__device__ __forceinline__ float quad_shr(float x, int q) {
    return dpp_f<0x110 + S>(q >= 4 - S ? 0.0f : x);
}

// The in-block prefixes of prefix16x2 with the 16 products held four per
// lane by a 4-lane row (lane q: products 4q and 4q + 2 as factors a * b,
// 4q + 1 and 4q + 3 as rounded values v): pair sums fma(a, b, v) (pairs 2q
// and 2q + 1 in lane q), their inclusive Kogge-Stone scan with offsets 1, 2,
// 4 (one pair within the lane and the last pair of lane q - 1; lane q - 1;
// lane q - 2), C_odd = P, C_even = fma(a, b, P of the pair before) -- the
// same operations on the same operands as the 8-lane form.
__device__ __forceinline__ void prefix16x4(float a0, float b0, float v1, float a2, float b2, float v3, float &c0,
                                           float &c1, float &c2, float &c3, int q) {
    const float q0 = __builtin_fmaf(a0, b0, v1), q1 = __builtin_fmaf(a2, b2, v3);
    float P0 = q0 + quad_shr<1>(q1, q);
    float P1 = q1 + q0;
    P0 = P0 + quad_shr<1>(P0, q);
    P1 = P1 + quad_shr<1>(P1, q);
    P0 = P0 + quad_shr<2>(P0, q);
    P1 = P1 + quad_shr<2>(P1, q);
    c0 = __builtin_fmaf(a0, b0, quad_shr<1>(P1, q));
    c1 = P0;
    c2 = __builtin_fmaf(a2, b2, P0);
    c3 = P1;
}

}  // namespace

namespace {

// Sum over the 64 lanes in a fixed butterfly order (same result in every
// lane, same order in every run).
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

}  // namespace

}  // namespace culda
