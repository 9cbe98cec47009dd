// mh.hip -- gfx950 kernels of the held-out evaluation and the learned
// Metropolis-Hastings proposal (no counterpart in the reference).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.h"
#include "kernels.h"
#include "philox.h"

namespace culda {

// ------------------------------------------------------------ held-out LL
// One wave per slot (evaluated tokens of one word).  The word's phi_k =
// (n_wk + beta) / (n_k + V beta) is staged in LDS in FP64 and
// Q = alpha * sum_k phi_k is reduced in a fixed order; each lane then takes
// one token: S = sum over the doc's fold-in theta entries of count * phi_k, and
// adds log((S + Q) / (n_d + K alpha)).  Slot partials are summed on the
// host in slot order.
__global__ __launch_bounds__(64) void heldout_kernel(HeldoutArgs a) {
    __shared__ double ph[kMaxK];
    const int lane = threadIdx.x;
    const EvalSlot s = uniform_copy(a.slots + blockIdx.x);
    const int K = a.k;
    const double vbeta = (double)a.num_words * a.beta;
    const int32_t *row = a.nwk + (int64_t)s.word * K;
    double q = 0.0;
    for (int k = lane; k < K; k += kWave) {
        const double p = ((double)row[k] + a.beta) / ((double)a.nk[k] + vbeta);
        ph[k] = p;
        q += p;
    }
    q = wave_sum(q) * a.alpha;
    wave_sync();
    const double kalpha = (double)K * a.alpha;
    double ll = 0.0;
    for (int64_t i = s.begin + lane; i < s.end; i += kWave) {
        const DocMeta m = a.meta[a.tokens[i]];
        double sum = 0.0;
        for (uint32_t j = 0; j < m.nnz; ++j) {
            const uint32_t e = a.theta[m.start + j];
            sum += (double)(e >> kThetaCountShift) * ph[e & kThetaTopicMask];
        }
        ll += log((sum + q) / ((double)m.nd + kalpha));
    }
    ll = wave_sum(ll);
    if (lane == 0) a.partial[blockIdx.x] = ll;
}

hipError_t launch_heldout(const HeldoutArgs &a, hipStream_t s) {
    if (a.num_slots <= 0) return hipSuccess;
    hipLaunchKernelGGL(heldout_kernel, dim3(a.num_slots), dim3(kWave), 0, s, a);
    return hipGetLastError();
}


// ======================================================= learned MH proposal
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// tanh-approximated GELU and its derivative, through the logistic
// function: 0.5 (1 + tanh(u)) = s = 1 / (1 + e^(-2u)), u = c (x + a x^3),
// so gelu = x s and gelu' = s + 2 x s (1 - s) c (1 + 3 a x^2) -- one
// hardware exp and one hardware reciprocal (v_rcp_f32, 1 ulp; 1 / inf = 0)
// instead of a tanh or a division.
constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kGeluA = 0.044715f;

__device__ __forceinline__ float gelu_logistic(float x) {
    const float u = kGeluC * (x + kGeluA * x * x * x);
    return __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * u));
}

__device__ __forceinline__ float gelu(float x) { return x * gelu_logistic(x); }

__device__ __forceinline__ float gelu_grad(float x) {
    const float sg = gelu_logistic(x);
    return sg + 2.0f * x * sg * (1.0f - sg) * kGeluC * (1.0f + 3.0f * kGeluA * x * x);
}

__device__ __forceinline__ float wave_sum_f(float x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

__device__ __forceinline__ float wave_max_f(float x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = fmaxf(x, __shfl_xor(x, d, kWave));
    return x;
}

// wave_*_all: the total of a 64-lane DPP scan (device.h) in every lane.
__device__ __forceinline__ uint32_t wave_add_all_u32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_add_scan_u32(x), kWave - 1);
}

__device__ __forceinline__ uint32_t wave_max_all_u32(uint32_t x) {
    x = wave_scan_u32(x, 0u, [](uint32_t a, uint32_t b) { return a > b ? a : b; });
    return (uint32_t)__builtin_amdgcn_readlane((int)x, kWave - 1);
}

__device__ __forceinline__ uint32_t wave_min_all_u32(uint32_t x) {
    x = wave_scan_u32(x, 0xffffffffu, [](uint32_t a, uint32_t b) { return a < b ? a : b; });
    return (uint32_t)__builtin_amdgcn_readlane((int)x, kWave - 1);
}

__device__ __forceinline__ float wave_max_all_f(float x) {
    const uint32_t ninf = 0xff800000u;
    uint32_t v = __float_as_uint(x);
    v = wave_scan_u32(v, ninf, [](uint32_t a, uint32_t b) {
        return __float_as_uint(fmaxf(__uint_as_float(a), __uint_as_float(b)));
    });
    return __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)v, kWave - 1));
}

__device__ __forceinline__ float wave_sum_all_f(float x) {
    uint32_t v = __float_as_uint(x);
    v = wave_scan_u32(v, 0u, [](uint32_t a, uint32_t b) {
        return __float_as_uint(__uint_as_float(a) + __uint_as_float(b));
    });
    return __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)v, kWave - 1));
}

// Signed count sketch of a doc's topic counts: topic k adds +-n_dk to
// bucket h(k); h and the sign are fixed multiplicative hashes of k.
__device__ __host__ __forceinline__ uint32_t sketch_bucket(uint32_t k) { return (k * 2654435761u) >> 24; }
__device__ __host__ __forceinline__ bool sketch_negative(uint32_t k) { return ((k * 0x85EBCA6Bu) >> 31) != 0; }

}  // namespace

// ------------------------------------------------------------ BF16 GEMM
// Both kernels take A stored [m][k] and B stored [n][k] (k-contiguous).
// 64 x 64 C tile per workgroup of four waves (2 x 2, each 32 x 32 = 2 x 2
// MFMA 16x16x32 tiles), for shapes that are not multiples of 128; A and B
// staged through LDS 32 deep as [row][k] so that every lane reads its 8
// consecutive k as one 16-byte LDS load.  MFMA lane layout (checked on
// gfx950): lane l feeds A[l % 16][8 (l / 16) .. +7] and
// B[8 (l / 16) .. +7][l % 16] and receives C[4 (l / 16) + r][l % 16].
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[kGemmTile][kGemmK + 8];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kGemmTile][kGemmK + 8];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t m0 = (int64_t)blockIdx.y * kGemmTile, n0 = (int64_t)blockIdx.x * kGemmTile;
    const int64_t kz = (int64_t)blockIdx.z * g.k;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.k; k0 += kGemmK) {
        {
            const int r = t >> 2, c8 = (t & 3) * 8;
            *reinterpret_cast<uint4 *>(&As[r][c8]) =
                *reinterpret_cast<const uint4 *>(g.a + (m0 + r) * g.lda + kz + k0 + c8);
            *reinterpret_cast<uint4 *>(&Bs[r][c8]) =
                *reinterpret_cast<const uint4 *>(g.b + (n0 + r) * g.ldb + kz + k0 + c8);
        }
        __syncthreads();
        bf16x8 af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            af[i] = *reinterpret_cast<const bf16x8 *>(&As[wm * 32 + i * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            bfr[j] = *reinterpret_cast<const bf16x8 *>(&Bs[wn * 32 + j * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
                const int64_t col = n0 + wn * 32 + j * 16 + (lane & 15);
                float x = acc[i][j][r];
                if (g.epilogue == kEpiGeluGrad) {
                    g.c16[row * g.ldc16 + col] = f2bf(x * gelu_grad(g.aux[row * g.ldaux + col]));
                } else {
                    if (g.bias) x += g.bias[col];
                    if (g.c) g.c[blockIdx.z * g.split_stride + row * g.ldc + col] = x;
                    if (g.c16) g.c16[row * g.ldc16 + col] = f2bf(g.epilogue == kEpiBiasGelu ? gelu(x) : x);
                }
            }
}

// Four FP32 values to BF16 (round to nearest even), packed for one 8-byte store.
__device__ __forceinline__ uint2 f2bf4(float a, float b, float c, float d) {
    return make_uint2((uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16), (uint32_t)f2bf(c) | ((uint32_t)f2bf(d) << 16));
}

// Epilogue of four consecutive columns col .. col + 3 of C row `row` (the
// four values of one MFMA lane), for the 128- and 256-row tile kernels.
__device__ __forceinline__ void gemm_epilogue4(const GemmArgs &g, int64_t row, int64_t col, float4 bias, float x0,
                                               float x1, float x2, float x3) {
    if (g.epilogue == kEpiGeluGrad) {
        const float4 p = *reinterpret_cast<const float4 *>(g.aux + row * g.ldaux + col);
        *reinterpret_cast<uint2 *>(g.c16 + row * g.ldc16 + col) =
            f2bf4(x0 * gelu_grad(p.x), x1 * gelu_grad(p.y), x2 * gelu_grad(p.z), x3 * gelu_grad(p.w));
        return;
    }
    x0 += bias.x;
    x1 += bias.y;
    x2 += bias.z;
    x3 += bias.w;
    if (g.c)
        *reinterpret_cast<float4 *>(g.c + blockIdx.z * g.split_stride + row * g.ldc + col) = make_float4(x0, x1, x2, x3);
    if (g.c16) {
        if (g.epilogue == kEpiBiasGelu) {
            x0 = gelu(x0);
            x1 = gelu(x1);
            x2 = gelu(x2);
            x3 = gelu(x3);
        }
        *reinterpret_cast<uint2 *>(g.c16 + row * g.ldc16 + col) = f2bf4(x0, x1, x2, x3);
    }
}

// 128 x 128 C tile per workgroup of four waves (2 x 2, each 64 x 64 =
// 4 x 4 MFMA 16x16x32 tiles) for A stored [m][k] and B stored [n][k]
// (both k-contiguous).  32-deep k stages, double-buffered in LDS: the next
// stage's global loads are issued before the current stage's MFMAs and
// stored to the other buffer after them.  The MFMA takes B's fragment as
// its first operand, so it forms C^T: lane l holds C[row l % 16][columns
// 4 (l / 16) .. +3] of each 16 x 16 tile, and the epilogue stores four
// consecutive columns at once (16 bytes FP32, 8 bytes BF16).  Tiles are
// dealt to the eight XCDs in contiguous runs (workgroup i runs on XCD
// i mod 8), so the column tiles of one row block share an L2.  Same
// epilogues as the 64-tile kernel.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void gemm_bf16_nt128_kernel(GemmArgs g) {
    constexpr int T = 128, KS = kGemmK, LD = KS + 8;
    __shared__ __attribute__((aligned(16))) uint16_t As[2][T][LD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][T][LD];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    int64_t tm = blockIdx.y, tn = blockIdx.x;
    {
        const uint32_t nt = gridDim.x, tiles = gridDim.x * gridDim.y;
        if (tiles % kXcds == 0) {
            const uint32_t id = blockIdx.y * nt + blockIdx.x;
            const uint32_t tile = (id % kXcds) * (tiles / kXcds) + id / kXcds;
            tm = tile / nt;
            tn = tile % nt;
        }
    }
    const int64_t m0 = tm * T, n0 = tn * T;
    const int64_t kz = (int64_t)blockIdx.z * g.k;
    const int lr = t >> 2, lc = (t & 3) * 8;  // rows lr and lr + 64, k offset lc
    const uint16_t *pa = g.a + (m0 + lr) * g.lda + kz + lc;
    const uint16_t *pb = g.b + (n0 + lr) * g.ldb + kz + lc;
    const int64_t a64 = 64 * g.lda, b64 = 64 * g.ldb;
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    uint4 ra0 = *reinterpret_cast<const uint4 *>(pa), ra1 = *reinterpret_cast<const uint4 *>(pa + a64);
    uint4 rb0 = *reinterpret_cast<const uint4 *>(pb), rb1 = *reinterpret_cast<const uint4 *>(pb + b64);
    *reinterpret_cast<uint4 *>(&As[0][lr][lc]) = ra0;
    *reinterpret_cast<uint4 *>(&As[0][lr + 64][lc]) = ra1;
    *reinterpret_cast<uint4 *>(&Bs[0][lr][lc]) = rb0;
    *reinterpret_cast<uint4 *>(&Bs[0][lr + 64][lc]) = rb1;
    __syncthreads();
    const int nk = g.k / KS;
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            const int64_t o = (int64_t)(kt + 1) * KS;
            ra0 = *reinterpret_cast<const uint4 *>(pa + o);
            ra1 = *reinterpret_cast<const uint4 *>(pa + a64 + o);
            rb0 = *reinterpret_cast<const uint4 *>(pb + o);
            rb1 = *reinterpret_cast<const uint4 *>(pb + b64 + o);
        }
        bf16x8 af[4], bfr[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            af[i] = *reinterpret_cast<const bf16x8 *>(&As[cur][wm * 64 + i * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            bfr[j] = *reinterpret_cast<const bf16x8 *>(&Bs[cur][wn * 64 + j * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j], af[i], acc[i][j], 0, 0, 0);
        if (more) {
            const int nx = cur ^ 1;
            *reinterpret_cast<uint4 *>(&As[nx][lr][lc]) = ra0;
            *reinterpret_cast<uint4 *>(&As[nx][lr + 64][lc]) = ra1;
            *reinterpret_cast<uint4 *>(&Bs[nx][lr][lc]) = rb0;
            *reinterpret_cast<uint4 *>(&Bs[nx][lr + 64][lc]) = rb1;
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int64_t col = n0 + wn * 64 + j * 16 + 4 * (lane >> 4);
        float4 bias = make_float4(0.f, 0.f, 0.f, 0.f);
        if (g.bias && g.epilogue != kEpiGeluGrad) bias = *reinterpret_cast<const float4 *>(g.bias + col);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t row = m0 + wm * 64 + i * 16 + (lane & 15);
            gemm_epilogue4(g, row, col, bias, acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]);
        }
    }
}

// Large products (the proposal queries: tens of thousands of rows): C tile
// 128 x 256 per workgroup of eight waves (2 x 4, each 64 x 64 = 4 x 4 MFMA
// 16x16x32 tiles), two workgroups per CU.  The operands stream through a
// ring of three LDS slots of 32-deep k slices filled by direct
// global-to-LDS loads (global_load_lds_dwordx4: each wave instruction writes
// 16 tile rows of 64 bytes, lane i the 16 bytes at i * 16), two slices in
// flight while the third is multiplied; one barrier per slice both publishes
// the slice and frees the slot read before it.  Tile rows (128 A rows, then
// 256 B rows) are 64 bytes, their four 16-byte k chunks XOR-swizzled by the
// row (ring_chunk) so that the fragment reads of 16 consecutive rows hit
// distinct LDS banks.  BF16 is the only output (kEpiStore or kEpiBiasGelu
// into c16; the launcher sends every other product to the 128-tile
// kernel): the epilogue stages the C tile through the idle ring (16-byte
// chunks swizzled by the row) and writes whole 512-byte rows.  Operand and
// fragment layouts and the XCD-aware tile order are the 128-tile kernel's.
typedef __attribute__((address_space(3))) void *lds_void_ptr;
typedef __attribute__((address_space(1))) void *global_void_ptr;
constexpr int kRingTM = 128, kRingTN = 256, kRingWaves = 8, kRingSlots = 3;

// physical 16-byte chunk of logical k chunk c (0..3) in tile row r
__device__ __forceinline__ int ring_chunk(int r, int c) { return c ^ ((r >> 1) & 3); }

// s_waitcnt immediate: wait until at most n vector memory operations of this
// wave are outstanding (other counters unconstrained)
#define CULDA_VMCNT(n) (0x0F70 | ((n)&15) | (((n) >> 4) << 14))

__global__ __launch_bounds__(64 * kRingWaves) __attribute__((amdgpu_waves_per_eu(4))) void gemm_bf16_ring_kernel(
    GemmArgs g) {
    constexpr int TM = kRingTM, TN = kRingTN, WN = 4, NW = kRingWaves, NB = kRingSlots, KS = kGemmK;
    constexpr int FM = TM / 2 / 16, FN = TN / WN / 16;  // 4 x 4 MFMA tiles per wave
    constexpr int ROWS = TM + TN, PER = ROWS / 16 / NW;  // 1 KB loads per wave per slice
    static_assert(KS == 32 && ROWS % (16 * NW) == 0 && TM * TN <= NB * ROWS * KS, "ring tile");
    __shared__ __attribute__((aligned(1024))) uint16_t ring[NB][ROWS * KS];
    const int t = threadIdx.x, lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wm = wave / WN, wn = wave % WN;
    int64_t tm = blockIdx.y, tn = blockIdx.x;
    {
        const uint32_t nt = gridDim.x, tiles = gridDim.x * gridDim.y;
        if (tiles % kXcds == 0) {
            const uint32_t id = blockIdx.y * nt + blockIdx.x;
            const uint32_t tile = (id % kXcds) * (tiles / kXcds) + id / kXcds;
            tm = tile / nt;
            tn = tile % nt;
        }
    }
    const int64_t m0 = tm * TM, n0 = tn * TN;
    // this lane's source of each of its wave's PER loads of a slice: load q
    // covers tile rows 16 j .. 16 j + 15, j = q * NW + wave
    const uint16_t *src[PER];
    {
        const int r16 = lane >> 2, c = ring_chunk(r16, lane & 3);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int row = (q * NW + wave) * 16 + r16;
            src[q] = row < TM ? g.a + (m0 + row) * g.lda + 8 * c : g.b + (n0 + row - TM) * g.ldb + 8 * c;
        }
    }
    auto issue = [&](int kt, int slot) {
#pragma unroll
        for (int q = 0; q < PER; ++q)
            __builtin_amdgcn_global_load_lds((global_void_ptr)(src[q] + (int64_t)kt * KS),
                                             (lds_void_ptr)(&ring[slot][(q * NW + wave) * 16 * KS]), 16, 0, 0);
    };
    // this lane's fragment: tile row (lane & 15) of a 16-row group, k chunk lane >> 4
    const int fo = (lane & 15) * KS + 8 * ring_chunk(lane & 15, lane >> 4);
    f32x4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nk = g.k / KS;
#pragma unroll
    for (int st = 0; st < NB - 1; ++st)
        if (st < nk) issue(st, st);
    for (int kt = 0; kt < nk; ++kt) {
        // slice kt landed for this wave (the next may stay in flight) ...
        if (kt + NB - 2 < nk)
            __builtin_amdgcn_s_waitcnt(CULDA_VMCNT(PER * (NB - 2)));
        else
            __builtin_amdgcn_s_waitcnt(CULDA_VMCNT(0));
        // ... and for every wave, which have also finished reading slice kt - 1
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        if (kt + NB - 1 < nk) issue(kt + NB - 1, (kt + NB - 1) % NB);
        const uint16_t *base = ring[kt % NB];
        bf16x8 bfr[FN];
#pragma unroll
        for (int j = 0; j < FN; ++j)
            bfr[j] = *reinterpret_cast<const bf16x8 *>(base + (TM + wn * (TN / WN) + j * 16) * KS + fo);
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const bf16x8 af = *reinterpret_cast<const bf16x8 *>(base + (wm * (TM / 2) + i * 16) * KS + fo);
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j], af, acc[i][j], 0, 0, 0);
        }
    }
    // staged BF16 epilogue: tile [TM][TN]
    constexpr int CPR = TN / 8, RPI = 64 / CPR;  // 16-byte chunks per row, rows per wave store
    uint16_t *tile = &ring[0][0];
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
#pragma unroll
    for (int j = 0; j < FN; ++j) {
        const int col = wn * (TN / WN) + j * 16 + 4 * (lane >> 4);
        float4 bias = make_float4(0.f, 0.f, 0.f, 0.f);
        if (g.bias) bias = *reinterpret_cast<const float4 *>(g.bias + n0 + col);
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const int row = wm * (TM / 2) + i * 16 + (lane & 15);
            float x0 = acc[i][j][0] + bias.x, x1 = acc[i][j][1] + bias.y, x2 = acc[i][j][2] + bias.z,
                  x3 = acc[i][j][3] + bias.w;
            if (g.epilogue == kEpiBiasGelu) {
                x0 = gelu(x0);
                x1 = gelu(x1);
                x2 = gelu(x2);
                x3 = gelu(x3);
            }
            const int ch = (col >> 3) ^ (row & (CPR - 1));
            *reinterpret_cast<uint2 *>(tile + row * TN + ch * 8 + (col & 4)) = f2bf4(x0, x1, x2, x3);
        }
    }
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
#pragma unroll
    for (int q = 0; q < TM / RPI / NW; ++q) {
        const int row = (q * NW + wave) * RPI + lane / CPR, ch = lane % CPR;
        const uint4 v = *reinterpret_cast<const uint4 *>(tile + row * TN + (ch ^ (row & (CPR - 1))) * 8);
        *reinterpret_cast<uint4 *>(g.c16 + (m0 + row) * g.ldc16 + n0 + ch * 8) = v;
    }
}

// BF16 transpose through 64 x 64 LDS tiles: dst[c][r] = src[r][c] for
// r < rows, c < cols (both multiples of 64).
__global__ __launch_bounds__(256) void transpose_bf16_kernel(const uint16_t *src, int64_t lds_, uint16_t *dst,
                                                             int64_t ldd) {
    __shared__ uint16_t tile[64][64 + 2];
    const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) tile[r][tx] = src[(r0 + r) * lds_ + c0 + tx];
    __syncthreads();
    for (int c = ty; c < 64; c += 4) dst[(c0 + c) * ldd + r0 + tx] = tile[tx][c];
}

// ------------------------------------------------------------ features
// One wave per doc of a unit: integer bucket sums in LDS (exact, so the
// result does not depend on the order), then bucket / n_d rounded to BF16.
__global__ __launch_bounds__(64) void sketch_kernel(const DocMeta *meta, const uint16_t *theta, int32_t num_docs,
                                                    uint16_t *out) {
    __shared__ int32_t acc[kSketch];
    const int lane = threadIdx.x;
    const int d = blockIdx.x;
    if (d >= num_docs) return;
    for (int b = lane; b < kSketch; b += kWave) acc[b] = 0;
    wave_sync();
    const DocMeta m = meta[d];
    for (uint32_t j = lane; j < m.nnz; j += kWave) {
        const uint32_t e = theta[m.start + j];
        const uint32_t k = e & kThetaTopicMask;
        const int32_t c = (int32_t)(e >> kThetaCountShift);
        atomicAdd(&acc[sketch_bucket(k)], sketch_negative(k) ? -c : c);
    }
    wave_sync();
    const float nd = (float)m.nd;
    for (int b = lane; b < kSketch; b += kWave)
        out[(int64_t)d * kSketch + b] = m.nd ? f2bf((float)acc[b] / nd) : (uint16_t)0;
}

// Training rows: position p of the epoch's order maps to triple
// tau = (mul * p + add) mod T (a bijection: mul is coprime to T), i.e.
// snapshot tau / N, token tau % N.  X row = [sketch; embedding] (BF16);
// rows past `rows` are zero with target -1.  One thread per 16-byte piece.
__global__ void gather_train_kernel(const uint16_t *sketches, const uint16_t *zs, const uint32_t *tok_doc,
                                    const uint32_t *tok_word, const uint16_t *emb16, uint64_t mul, uint64_t add,
                                    uint64_t T, uint64_t N, int64_t D, int64_t p0, int32_t rows, int32_t rows_pad,
                                    uint16_t *X, int32_t *tgt, int32_t *wrow) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int kPieces = kIn / 8;
    const int64_t r = idx / kPieces;
    const int pc = (int)(idx % kPieces);
    if (r >= rows_pad) return;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (r < rows) {
        const uint64_t tau = (mul * (uint64_t)(p0 + r) + add) % T;
        const uint64_t slot = tau / N, i = tau % N;
        const uint32_t d = tok_doc[i], w = tok_word[i];
        if (pc < kSketch / 8)
            v = *reinterpret_cast<const uint4 *>(sketches + ((int64_t)slot * D + d) * kSketch + pc * 8);
        else
            v = *reinterpret_cast<const uint4 *>(emb16 + (int64_t)w * kEmb + (pc - kSketch / 8) * 8);
        if (pc == 0) {
            tgt[r] = zs[slot * N + i];
            wrow[r] = (int32_t)w;
        }
    } else if (pc == 0) {
        tgt[r] = -1;
        wrow[r] = -1;
    }
    *reinterpret_cast<uint4 *>(X + r * kIn + pc * 8) = v;
}

// Query rows for explicit (global doc, word) pairs; rows past n are zero.
__global__ void gather_query_kernel(const uint16_t *sketch, const uint32_t *docs, const uint32_t *words,
                                    const uint16_t *emb16, int64_t n, int64_t rows_pad, uint16_t *X) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int kPieces = kIn / 8;
    const int64_t r = idx / kPieces;
    const int pc = (int)(idx % kPieces);
    if (r >= rows_pad) return;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (r < n) {
        if (pc < kSketch / 8)
            v = *reinterpret_cast<const uint4 *>(sketch + (int64_t)docs[r] * kSketch + pc * 8);
        else
            v = *reinterpret_cast<const uint4 *>(emb16 + (int64_t)words[r] * kEmb + (pc - kSketch / 8) * 8);
    }
    *reinterpret_cast<uint4 *>(X + r * kIn + pc * 8) = v;
}

// ------------------------------------------------------------ training
// One wave per row: softmax over K logits, loss = log sum exp - logit[z],
// G = (softmax - onehot(z)) * inv_rows in BF16 (zero for padding rows).
__global__ __launch_bounds__(64) void softmax_ce_kernel(const float *logits, int32_t K, const int32_t *tgt,
                                                        float inv_rows, uint16_t *G, float *loss) {
    const int lane = threadIdx.x;
    const int64_t r = blockIdx.x;
    const float *l = logits + r * K;
    uint16_t *gr = G + r * K;
    const int32_t z = tgt[r];
    if (z < 0) {
        for (int k = lane; k < K; k += kWave) gr[k] = 0;
        if (lane == 0) loss[r] = 0.f;
        return;
    }
    float mx = -INFINITY;
    for (int k = lane; k < K; k += kWave) mx = fmaxf(mx, l[k]);
    mx = wave_max_f(mx);
    float s = 0.f;
    for (int k = lane; k < K; k += kWave) s += expf(l[k] - mx);
    s = wave_sum_f(s);
    const float inv_s = 1.0f / s;
    for (int k = lane; k < K; k += kWave) {
        const float p = expf(l[k] - mx) * inv_s;
        gr[k] = f2bf((p - (k == z ? 1.0f : 0.0f)) * inv_rows);
    }
    if (lane == 0) loss[r] = logf(s) + mx - l[z];
}

// Row losses summed per block of 256 in a fixed tree order, each block's
// sum added in 2^-24 units with one integer atomic (exact, order-free).
__global__ __launch_bounds__(256) void loss_sum_kernel(const float *loss, int32_t n, unsigned long long *loss_fx) {
    __shared__ float sh[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    sh[threadIdx.x] = i < n ? loss[i] : 0.f;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss_fx, (unsigned long long)(long long)rint((double)sh[0] * 16777216.0));
}

// Column sums of a BF16 matrix (bias gradients) in two fixed-order
// stages: block (c, p) sums rows [p * kColRows, ...) of its 256 columns into
// part[p][col]; then each column adds its parts in order.
constexpr int kColRows = 128;

__global__ void colsum_part_kernel(const uint16_t *x, int64_t ld, int32_t rows, int32_t cols, float *part) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    const int r0 = blockIdx.y * kColRows, r1 = min(rows, r0 + kColRows);
    float s = 0.f;
#pragma unroll 8
    for (int r = r0; r < r1; ++r) s += __uint_as_float((uint32_t)x[(int64_t)r * ld + j] << 16);
    part[(int64_t)blockIdx.y * cols + j] = s;
}

// out[j] = sum over p < parts of part[p][j], in order (split-K and column-sum reductions).
__global__ void sum_parts_kernel(const float *part, int32_t parts, int64_t n, float *out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float s = 0.f;
#pragma unroll 8
    for (int p = 0; p < parts; ++p) s += part[p * n + j];
    out[j] = s;
}

// Embedding gradient: row r adds dX_emb[r] to word wrow[r]'s row, in
// fixed point (2^-32 units, int64 atomics), so the sum is exact and the
// same in every run whatever the order of the additions.
__global__ void emb_grad_kernel(const float *dxe, const int32_t *wrow, int32_t rows, long long *acc) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)rows * kEmb) return;
    const int64_t r = idx / kEmb;
    const int c = (int)(idx % kEmb);
    const int32_t w = wrow[r];
    if (w < 0) return;
    const long long q = (long long)rint((double)dxe[idx] * 4294967296.0);
    if (q) atomicAdd(reinterpret_cast<unsigned long long *>(acc + (int64_t)w * kEmb + c), (unsigned long long)q);
}

// Adam (torch.optim.Adam's update order): m = b1 m + (1 - b1) g;
// v = b2 v + (1 - b2) g^2; p += -step * m / (sqrt(v) / sqrt_bc2 + eps).
// The gradient is FP32 (g) or fixed point (g64, cleared here).  Also
// refreshes the BF16 copy.
__global__ void adam_kernel(float *p, float *m, float *v, const float *g, long long *g64, uint16_t *p16,
                            uint16_t *p16t, int64_t rows, int64_t n, float step, float b1, float b2, float eps,
                            float sqrt_bc2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi;
    if (g) {
        gi = g[i];
    } else {
        gi = (float)((double)g64[i] * (1.0 / 4294967296.0));
        g64[i] = 0;
    }
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] + (-step) * (mi / (sqrtf(vi) / sqrt_bc2 + eps));
    p[i] = pi;
    if (p16) p16[i] = f2bf(pi);
    if (p16t) {  // [cols][rows] copy of a [rows][cols] matrix
        const int64_t cols = n / rows, r = i / cols, c = i % cols;
        p16t[c * rows + r] = f2bf(pi);
    }
}

__global__ void to_bf16_kernel(const float *x, uint16_t *y, uint16_t *yt, int64_t rows, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint16_t h = f2bf(x[i]);
    y[i] = h;
    if (yt) {
        const int64_t cols = n / rows;
        yt[(i % cols) * rows + i / cols] = h;
    }
}

// ------------------------------------------------------------ MH step
// Alias-table entry of outcome k with weight Wk and rank rk among the
// lights (Wk < M) or heavies, from the prefix sums in sDE (lights' D at
// [0, nL), heavies' E at [nL, nL + nH)) and the heavies' ids in sHI:
// light: F = Wk, alias = first heavy b with E_b > D_{rk-1};
// last heavy: F = M, alias = itself; other heavy b = rk: emptied after
// a* = min{a : D_a >= E_b} lights (0 if E_b = 0), F = M + E_b - D_{a*},
// alias = heavy b + 1.
__device__ __forceinline__ void alias_entry(uint32_t k, uint32_t Wk, uint32_t rk, const uint32_t *sDE,
                                            const uint16_t *sHI, uint32_t nL, uint32_t nH, uint32_t &F,
                                            uint32_t &A) {
    constexpr uint32_t M = 1u << kAliasBits;
    if (Wk < M) {
        const uint32_t dprev = rk ? sDE[rk - 1] : 0u;
        uint32_t lo = 0, hi = nH;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sDE[nL + mid] > dprev) hi = mid; else lo = mid + 1;
        }
        F = Wk;
        A = sHI[lo < nH ? lo : nH - 1];
    } else if (rk + 1 == nH) {
        F = M;
        A = k;
    } else {
        const uint32_t eb = sDE[nL + rk];
        uint32_t astar = 0;
        if (eb) {
            uint32_t lo = 0, hi = nL;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sDE[mid] >= eb) hi = mid; else lo = mid + 1;
            }
            astar = lo + 1;
        }
        F = M + eb - (astar ? sDE[astar - 1] : 0u);
        A = sHI[rk + 1];
    }
}

// One wave per token.  Proposal weights: W_k = max(1, floor(softmax_k *
// K * 2^20)), the remainder added to the largest weight so that
// sum W = K * 2^20 exactly; the proposal is then exactly W_k / (K 2^20).
// Its Walker alias table (integer sweep: lights and heavies in index order
// with inclusive prefix deficits D and excesses E, see alias_entry) keeps
// every weight's mass exactly.  Only the prefix sums are built; the entry
// of each drawn bucket is evaluated from them.  Then `proposals`
// independence-MH steps on the token's collapsed conditional with frozen
// counts (own assignment included, as in the plain sampler),
// p(k) = (n_dk + alpha)(n_wk + beta) / (n_k + V beta) in FP64: propose k'
// from the table (bucket j = K u0, threshold test on 20 more bits), accept
// iff u (p(k) W_k') < p(k') W_k.  Draws: Philox((g, sweep, 2, r), (seed, 0)).
__global__ __launch_bounds__(64) void mh_kernel(MHArgs a) {
    // sDE: light prefix deficits at [0, nL), heavy prefix excesses at [nL, K)
    __shared__ uint32_t sW[kMaxK], sDE[kMaxK];
    __shared__ uint16_t sHI[kMaxK];
    const int lane = threadIdx.x;
    const int r = blockIdx.x;
    if (r >= a.n) return;
    const int K = a.model.k, KL = K / kWave;
    constexpr uint32_t M = 1u << kAliasBits;
    // the token, and its doc's theta row (<= kMHThetaEntries entries, two
    // per dword, 16 dwords per lane; unused entries hold 0 = topic 0 with
    // count 0), loaded first
    const int64_t i = a.start + r;
    const uint32_t d = a.unit.tokens[i];
    const uint32_t w = a.word[i];
    const uint32_t z0 = a.unit.z[i];
    const uint64_t gid = (uint64_t)(i + a.gdelta[w]);
    const DocMeta dm = a.unit.meta[d];
    uint32_t tr[16];
    const uint32_t *trow = reinterpret_cast<const uint32_t *>(a.unit.theta + dm.start);
    // dwords of the row per lane actually holding entries (128 entries per
    // dword index t); the rest stay 0 and are skipped below
    const int tmax = (int)((dm.nnz + 127u) >> 7);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const uint32_t j = lane + kWave * t;  // entries 2j, 2j + 1 (2j + 1 past nnz is a zero of the slab tail)
        tr[t] = (t < tmax && 2 * j < dm.nnz) ? trow[j] : 0u;
    }
    const uint16_t *lg = a.logits + (int64_t)r * K + lane * KL;
    float e[16];
    float mx = -INFINITY;
    if (KL == 16) {
        // the lane's 16 logits (32 bytes, aligned): two 16-byte loads
        const u32x4 l0 = load_global4(lg), l1 = load_global4(lg + 8);
        const uint32_t lw[8] = {l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, l1.z, l1.w};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            e[2 * j] = __uint_as_float(lw[j] << 16);
            e[2 * j + 1] = __uint_as_float(lw[j] & 0xffff0000u);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) mx = fmaxf(mx, e[j]);
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            e[j] = j < KL ? __uint_as_float(load_global_u16(lg + j) << 16) : -INFINITY;
            mx = fmaxf(mx, e[j]);
        }
    }
    mx = wave_max_all_f(mx);
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        e[j] = j < KL ? __expf(e[j] - mx) : 0.f;
        s += e[j];
    }
    s = wave_sum_all_f(s);
    const uint32_t KM = (uint32_t)K << kAliasBits;
    const float scale = (float)KM / s;
    uint32_t W[16];
    uint32_t lsum = 0, lmax = 0;  // lsum wraps mod 2^32 (see rem below)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if (j < KL) {
            const float x = e[j] * scale;
            uint32_t wv = x >= 1.0f ? (uint32_t)x : 1u;
            if (wv > KM) wv = KM;
            W[j] = wv;
            lsum += wv;
            lmax = wv > lmax ? wv : lmax;
        } else {
            W[j] = 0;
        }
    }
    // the sum of all weights is within a few K of K 2^20 < 2^31, so the
    // remainder is exact from sums modulo 2^32
    const uint32_t wsum = (uint32_t)__builtin_amdgcn_readlane((int)wave_add_scan_u32(lsum), kWave - 1);
    const int32_t rem = (int32_t)(KM - wsum);
    // the largest weight, lowest topic among equals, takes the remainder
    const uint32_t wmax = wave_max_all_u32(lmax);
    uint32_t lk = 0xffffffffu;
#pragma unroll
    for (int j = 15; j >= 0; --j)
        if (j < KL && W[j] == wmax) lk = lane * KL + j;
    const uint32_t kmax = wave_min_all_u32(lk);
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (j < KL && (uint32_t)(lane * KL + j) == kmax) W[j] = (uint32_t)((int32_t)W[j] + rem);
    // lane-local light / heavy counts and sums, then wave prefixes
    uint32_t cL = 0, sDef = 0, cH = 0, sExc = 0;
    uint32_t lmask = 0;  // bit j: element j of this lane is a light
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (j < KL) {
            if (W[j] < M) {
                ++cL;
                sDef += M - W[j];
                lmask |= 1u << j;
            } else {
                ++cH;
                sExc += W[j] - M;
            }
        }
    // exclusive prefixes (inclusive DPP scan minus the lane's own value)
    const uint32_t iL = wave_add_scan_u32(cL), iD = wave_add_scan_u32(sDef);
    const uint32_t iH = wave_add_scan_u32(cH), iE = wave_add_scan_u32(sExc);
    const uint32_t nL = (uint32_t)__builtin_amdgcn_readlane((int)iL, kWave - 1);
    const uint32_t nH = (uint32_t)__builtin_amdgcn_readlane((int)iH, kWave - 1);
    uint32_t aL = iL - cL, dAcc = iD - sDef, bH = iH - cH, eAcc = iE - sExc;
    const uint32_t aL0 = aL, bH0 = bH;  // ranks of this lane's first light / heavy
    // rank of this lane's element jj among the lights or the heavies
    auto rank_of = [&](uint32_t jj) {
        uint32_t wj = 0, nl = 0, nh = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            if ((uint32_t)q == jj) wj = W[q];
            if ((uint32_t)q < jj && q < KL) {
                if (W[q] < M) ++nl; else ++nh;
            }
        }
        return wj < M ? aL0 + nl : bH0 + nh;
    };
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (j < KL) {
            const uint32_t k = lane * KL + j;
            sW[k] = W[j];
            if (W[j] < M) {
                dAcc += M - W[j];
                sDE[aL++] = dAcc;
            } else {
                eAcc += W[j] - M;
                sDE[nL + bH] = eAcc;
                sHI[bH++] = (uint16_t)k;
            }
        }
    wave_sync();
    if (a.w_out)  // whole tables, for inspection
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (j < KL) {
                const uint32_t k = lane * KL + j;
                uint32_t f, al;
                alias_entry(k, W[j], rank_of((uint32_t)j), sDE, sHI, nL, nH, f, al);
                a.w_out[(int64_t)r * K + k] = (int32_t)W[j];
                a.f_out[(int64_t)r * K + k] = f;
                a.a_out[(int64_t)r * K + k] = al;
            }
    // MH steps (wave-uniform): every candidate (current topic, then each
    // proposal) is drawn first -- the alias entry of a drawn bucket is
    // evaluated from the prefix sums -- its n_dk comes from the doc's theta
    // row held in registers (the counts of its entries, a wave sum), its
    // n_wk / n_k from memory, all loads issued together.
    const int32_t *nrow = a.model.nwk + (int64_t)w * K;
    const double alpha = a.model.alpha, beta = a.model.beta;
    const double vbeta = (double)a.model.num_words * beta;
    const int np = a.proposals;
    uint32_t cand[kMaxProposals + 1];
    double ud[kMaxProposals];
    cand[0] = z0;
    // every proposal's Philox block in one evaluation: lane p computes
    // proposal p's (the counters differ only in p)
    const u64x4 xl = philox4x64_10(gid, a.sweep, kPurposeMH, (uint64_t)(lane < np ? lane : 0), a.model.seed, 0);
#pragma unroll
    for (int p = 0; p < kMaxProposals; ++p) {
        if (p < np) {
            const uint64_t x0 = (uint64_t)readlane64((int64_t)xl.v[0], p), x1 = (uint64_t)readlane64((int64_t)xl.v[1], p);
            const uint64_t x2 = (uint64_t)readlane64((int64_t)xl.v[2], p);
            const uint32_t j = (uint32_t)mulhi64(x0, (uint64_t)K);
            const uint32_t rr = (uint32_t)(x1 >> (64 - kAliasBits));
            ud[p] = (double)(x2 >> 11) * 0x1.0p-53;
            // the rank of bucket j among the lights or the heavies, from the
            // lane that holds it: its first rank plus the lights (heavies)
            // before element jj of that lane
            const uint32_t wj = sW[j], jj = j % (uint32_t)KL;
            const uint32_t below = (1u << jj) - 1u;
            const uint32_t own = wj < M ? aL0 + (uint32_t)__builtin_popcount(lmask & below)
                                        : bH0 + (uint32_t)__builtin_popcount(~lmask & below);
            const uint32_t rk = (uint32_t)__shfl((int)own, (int)(j / (uint32_t)KL), kWave);
            uint32_t f, al;
            alias_entry(j, wj, rk, sDE, sHI, nL, nH, f, al);
            cand[p + 1] = rr < f ? j : al;
        } else {
            ud[p] = 0.0;
            cand[p + 1] = z0;
        }
    }
    double pt[kMaxProposals + 1];
#pragma unroll
    for (int c = 0; c <= kMaxProposals; ++c) {
        if (c > np) {
            pt[c] = 0.0;
            continue;
        }
        const uint32_t k = cand[c];
        // n_dk = sum of the counts of the topic's entries (a padding 0 adds 0)
        uint32_t cnt = 0;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            if (t >= tmax) break;  // wave-uniform
            if ((tr[t] & kThetaTopicMask) == k) cnt += (tr[t] >> kThetaCountShift) & kThetaMaxCount;
            if (((tr[t] >> 16) & kThetaTopicMask) == k) cnt += tr[t] >> (16 + kThetaCountShift);
        }
        cnt = wave_add_all_u32(cnt);
        pt[c] = (((double)cnt + alpha) * ((double)nrow[k] + beta)) / ((double)a.model.nk[k] + vbeta);
    }
    uint32_t cur = z0;
    double pc = pt[0];
    uint32_t acc = 0;
#pragma unroll
    for (int p = 0; p < kMaxProposals; ++p) {
        if (p < np) {
            const uint32_t kp = cand[p + 1];
            if (kp != cur && ud[p] * (pc * (double)sW[kp]) < pt[p + 1] * (double)sW[cur]) {
                cur = kp;
                pc = pt[p + 1];
                ++acc;
            }
        }
    }
    if (lane == 0) {
        a.unit.z[i] = (uint16_t)cur;
        if (acc) atomicAdd(a.accepted, (unsigned long long)acc);
    }
}

// ------------------------------------------------------------ launchers
hipError_t launch_gemm_bf16(const GemmArgs &g, hipStream_t s, int32_t splits) {
    if (g.m <= 0 || g.n <= 0 || g.k <= 0) return hipSuccess;
    if (splits < 1 || splits > 65535) return hipErrorInvalidValue;
    if (g.m % kGemmTile || g.n % kGemmTile || g.k % kGemmK || g.lda % 8 || g.ldb % 8 ||
        ((uintptr_t)g.a & 15) || ((uintptr_t)g.b & 15) || g.m / kGemmTile > 65535)
        return hipErrorInvalidValue;
    if (g.epilogue == kEpiGeluGrad && (!g.c16 || !g.aux)) return hipErrorInvalidValue;
    // the 128-tile kernel stores four consecutive columns at once: 16-byte
    // aligned FP32 rows and bias, 8-byte aligned BF16 rows
    const bool vec = g.ldc % 4 == 0 && g.ldc16 % 4 == 0 && g.ldaux % 4 == 0 && g.split_stride % 4 == 0 &&
                     ((uintptr_t)g.c & 15) == 0 && ((uintptr_t)g.aux & 15) == 0 && ((uintptr_t)g.bias & 15) == 0 &&
                     ((uintptr_t)g.c16 & 7) == 0;
    // large products with BF16 the only output (the proposal queries: at
    // least two tiles per CU): the ring kernel
    const bool bf16_only = g.c16 && !g.c && (g.epilogue == kEpiStore || g.epilogue == kEpiBiasGelu) &&
                           g.ldc16 % 8 == 0 && ((uintptr_t)g.c16 & 15) == 0;
    if (bf16_only && splits == 1 && g.m % kRingTM == 0 && g.n % kRingTN == 0 && vec &&
        (int64_t)(g.m / kRingTM) * (g.n / kRingTN) >= 512) {
        hipLaunchKernelGGL(gemm_bf16_ring_kernel, dim3(g.n / kRingTN, g.m / kRingTM), dim3(64 * kRingWaves), 0, s, g);
    } else if (g.m % 128 == 0 && g.n % 128 == 0 && vec)
        hipLaunchKernelGGL(gemm_bf16_nt128_kernel, dim3(g.n / 128, g.m / 128, splits), dim3(256), 0, s, g);
    else
        hipLaunchKernelGGL(gemm_bf16_kernel, dim3(g.n / kGemmTile, g.m / kGemmTile, splits), dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t launch_transpose_bf16(const uint16_t *src, int64_t rows, int64_t cols, int64_t ld_src, uint16_t *dst,
                                 int64_t ld_dst, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return hipSuccess;
    if (rows % 64 || cols % 64 || rows / 64 > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(transpose_bf16_kernel, dim3((unsigned)(cols / 64), (unsigned)(rows / 64)), dim3(256), 0, s, src,
                       ld_src, dst, ld_dst);
    return hipGetLastError();
}

hipError_t launch_sketch(const DocMeta *meta, const uint16_t *theta, int32_t num_docs, uint16_t *out,
                         hipStream_t s) {
    if (num_docs <= 0) return hipSuccess;
    hipLaunchKernelGGL(sketch_kernel, dim3(num_docs), dim3(kWave), 0, s, meta, theta, num_docs, out);
    return hipGetLastError();
}

hipError_t launch_gather_train(const uint16_t *sketches, const uint16_t *zs, const uint32_t *tok_doc,
                               const uint32_t *tok_word, const uint16_t *emb16, uint64_t mul, uint64_t add, uint64_t T,
                               uint64_t N, int64_t D, int64_t p0, int32_t rows, int32_t rows_pad, uint16_t *X,
                               int32_t *tgt, int32_t *wrow, hipStream_t s) {
    const int64_t n = (int64_t)rows_pad * (kIn / 8);
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_train_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sketches, zs, tok_doc,
                       tok_word, emb16, mul, add, T, N, D, p0, rows, rows_pad, X, tgt, wrow);
    return hipGetLastError();
}

hipError_t launch_gather_query(const uint16_t *sketch, const uint32_t *docs, const uint32_t *words,
                               const uint16_t *emb16, int64_t n, int64_t rows_pad, uint16_t *X, hipStream_t s) {
    const int64_t m = rows_pad * (kIn / 8);
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_query_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, sketch, docs, words,
                       emb16, n, rows_pad, X);
    return hipGetLastError();
}

hipError_t launch_softmax_ce(const float *logits, int32_t K, const int32_t *tgt, float inv_rows, uint16_t *G,
                             float *loss, unsigned long long *loss_fx, int32_t rows_pad, hipStream_t s) {
    if (rows_pad <= 0) return hipSuccess;
    hipLaunchKernelGGL(softmax_ce_kernel, dim3(rows_pad), dim3(kWave), 0, s, logits, K, tgt, inv_rows, G, loss);
    if (loss_fx)
        hipLaunchKernelGGL(loss_sum_kernel, dim3((rows_pad + 255) / 256), dim3(256), 0, s, loss, rows_pad, loss_fx);
    return hipGetLastError();
}

int32_t colsum_parts(int32_t rows) { return std::max(1, (rows + kColRows - 1) / kColRows); }

hipError_t launch_colsum_bf16(const uint16_t *x, int64_t ld, int32_t rows, int32_t cols, float *part, float *out,
                              hipStream_t s) {
    if (cols <= 0) return hipSuccess;
    const int32_t parts = colsum_parts(rows);
    hipLaunchKernelGGL(colsum_part_kernel, dim3((cols + 255) / 256, parts), dim3(256), 0, s, x, ld, rows, cols, part);
    hipLaunchKernelGGL(sum_parts_kernel, dim3((cols + 255) / 256), dim3(256), 0, s, part, parts, (int64_t)cols, out);
    return hipGetLastError();
}

// Deterministic split-K: split p computes the product over its share of k
// into part + p * m * n; the parts are then added in order into g.c.
hipError_t launch_gemm_bf16_splitk(const GemmArgs &g, int32_t splits, float *part, hipStream_t s) {
    if (splits <= 1) return launch_gemm_bf16(g, s);
    if (g.epilogue != kEpiStore || g.bias || g.c16 || !g.c || g.ldc != g.n || g.k % (splits * kGemmK))
        return hipErrorInvalidValue;
    GemmArgs q = g;
    q.k = g.k / splits;
    q.c = part;
    q.split_stride = (int64_t)g.m * g.n;
    const hipError_t e = launch_gemm_bf16(q, s, splits);
    if (e != hipSuccess) return e;
    const int64_t n = (int64_t)g.m * g.n;
    hipLaunchKernelGGL(sum_parts_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, splits, n, g.c);
    return hipGetLastError();
}

hipError_t launch_emb_grad(const float *dxe, const int32_t *wrow, int32_t rows, long long *acc, hipStream_t s) {
    const int64_t n = (int64_t)rows * kEmb;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(emb_grad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dxe, wrow, rows, acc);
    return hipGetLastError();
}

hipError_t launch_adam(float *p, float *m, float *v, const float *g, long long *g64, uint16_t *p16, uint16_t *p16t,
                       int64_t rows, int64_t n, float step, float b1, float b2, float eps, float sqrt_bc2,
                       hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (p16t && (rows <= 0 || n % rows)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, m, v, g, g64, p16, p16t,
                       rows, n, step, b1, b2, eps, sqrt_bc2);
    return hipGetLastError();
}

hipError_t launch_to_bf16(const float *x, uint16_t *y, uint16_t *yt, int64_t rows, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (yt && (rows <= 0 || n % rows)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, yt, rows, n);
    return hipGetLastError();
}

hipError_t launch_mh(const MHArgs &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    if (a.model.k % kWave || a.model.k > kMaxK || a.proposals < 0 || a.proposals > kMaxProposals)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(mh_kernel, dim3(a.n), dim3(kWave), 0, s, a);
    return hipGetLastError();
}

}  // namespace culda
