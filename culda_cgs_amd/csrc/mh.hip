// mh.hip -- gfx950 kernels of the held-out evaluation and the learned
// Metropolis-Hastings proposal (no counterpart in the reference).
#include <hip/hip_runtime.h>

#include "device.h"
#include "kernels.h"

namespace culda {

// ------------------------------------------------------------ held-out LL
// One wave per slot (evaluated tokens of one word).  The word's phi_k =
// (n_wk + beta) / (n_k + V beta) is staged in LDS in FP64 and
// Q = alpha * sum_k phi_k is reduced in a fixed order; each lane then takes
// one token: S = sum over the doc's fold-in nonzeros of n_dk * phi_k, and
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
    const uint32_t cb = a.count_bits, cmask = (1u << cb) - 1u;
    const double kalpha = (double)K * a.alpha;
    double ll = 0.0;
    for (int64_t i = s.begin + lane; i < s.end; i += kWave) {
        const DocMeta m = a.meta[a.tokens[i]];
        double sum = 0.0;
        for (uint32_t j = 0; j < m.nnz; ++j) {
            const uint32_t e = a.theta[m.start + j];
            sum += (double)(e & cmask) * ph[(e >> cb) >> 2];
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

}  // namespace culda
