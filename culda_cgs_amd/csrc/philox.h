// philox.h -- counter-based Philox4x64-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), host and device.
//
// Replaces the clock-seeded, shared curandState of the reference
// (kernel/lda_train_kernel.cu:14-18, :141): every draw is a pure function
// of (seed, global token id, sweep index, purpose), so z does not depend on
// the launch shape, the chunking or the number of devices.
//
// The stream is the one numpy.random.Philox generates: numpy's
// Philox(counter=c, key=k).random_raw(4) equals philox4x64_10(c + 1, k).
#pragma once
#include <stdint.h>

#if defined(__HIP__)
#define CULDA_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define CULDA_HD inline
#endif

namespace culda {

struct u64x4 {
    uint64_t v[4];
};

CULDA_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// 10 rounds; key schedule bumps the key between rounds.
CULDA_HD u64x4 philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0,
                             uint64_t k1) {
    const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
    const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += W0;
            k1 += W1;
        }
        const uint64_t hi0 = mulhi64(M0, c0), lo0 = M0 * c0;
        const uint64_t hi1 = mulhi64(M1, c2), lo1 = M1 * c2;
        const uint64_t n0 = hi1 ^ c1 ^ k0;
        const uint64_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    u64x4 out;
    out.v[0] = c0;
    out.v[1] = c1;
    out.v[2] = c2;
    out.v[3] = c3;
    return out;
}

// Purposes (third counter word).
// Counter word 2 separates the streams: z init, plain sweeps, and the
// learned-MH proposals (word 3 = proposal index within the sweep).
enum : uint64_t { kPurposeInit = 0, kPurposeSweep = 1, kPurposeMH = 2 };

// First output word for (global token id, sweep, purpose).
CULDA_HD uint64_t token_bits(uint64_t seed, uint64_t gid, uint64_t sweep, uint64_t purpose) {
    return philox4x64_10(gid, sweep, purpose, 0, seed, 0).v[0];
}

// Sweep draws: token g takes word g & 3 of Philox((g >> 2, sweep,
// kPurposeSweep, 0), (seed, 0)), so one block serves four tokens; the
// uniform on (0, 1] is (top 24 bits + 1) * 2^-24, exact in FP32.
CULDA_HD uint32_t uniform_bits(uint64_t x) { return (uint32_t)(x >> 40); }
CULDA_HD float bits24_to_uniform(uint32_t b) { return (float)(b + 1u) * 5.9604644775390625e-8f; }

// Initial topic: floor(x * k / 2^64).
CULDA_HD uint32_t bits_to_topic(uint64_t x, uint32_t k) { return (uint32_t)mulhi64(x, k); }

}  // namespace culda
