// kernels.hip -- gfx950 kernels of one plain CGS iteration.
//
// Built with -ffp-contract=off: every FP32 operation below is a separate,
// correctly rounded IEEE operation in the order written, so the CPU oracle
// can replay it bit for bit.  The order is the normative one:
//   phi_k   = fl(fl(fl(n_wk) + beta) / den_k),  den_k = fl(fl(n_k) + fl(beta * fl(V)))
//   Q bucket products a_k * phi_k with a_k = alpha; S bucket products
//   count_j * phi_{topic_j} (count_j exact in FP32);
//   blocks of 16 products; pair sums q_m = fma(x_2m, y_2m, fl(x_2m+1 *
//   y_2m+1)) (the even product fused, one rounding); a block total T_b is
//   the balanced tree of its 8 pair sums; S: block totals carried
//   sequentially R_{b+1} = fl(R_b + T_b); Q: 64-lane Kogge-Stone scan of
//   the block totals.  In-block prefixes: the pair sums' Kogge-Stone scan
//   P_m (offsets 1, 2, 4), prefix_2m+1 = P_m, prefix_2m = fma(x_2m, y_2m,
//   P_m-1) (P_-1 = 0).  t = fl(u * fl(S + Q)); S bucket iff t < S (target t), else
//   Q bucket (target fl(t - S)).  Search: first block whose inclusive
//   prefix exceeds the target, then first element of that block whose
//   fl(prefix_before_block + in-block prefix) exceeds it; fallbacks take
//   the last element.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "device.h"
#include "kernels.h"
#include "philox.h"

namespace culda {


// ------------------------------------------------------------ z init
// One wave per slot; z = mulhi64(Philox((g, 0, 0, 0), (seed, 0))[0], K).
__global__ __launch_bounds__(256) void init_z_kernel(InitArgs a) {
    const int slot = blockIdx.x * kWavesPerGroup + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (slot >= a.num_slots) return;
    const SlotDesc s = a.slots[slot];
    const UnitDev &u = a.units[s.unit];
    const ModelDev &m = a.models[u.model];
    for (int64_t i = s.begin + lane; i < s.end; i += kWave) {
        const uint64_t x = token_bits(m.seed, (uint64_t)(i + s.gdelta), 0, kPurposeInit);
        u.z[i] = (uint16_t)bits_to_topic(x, (uint32_t)m.k);
    }
}

__global__ void fill_u32_kernel(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void add_u32_kernel(uint32_t *p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] += v;
}

// Clear before a count rebuild: fill[0, n_fill) = 0 (the n_k vectors), and
// for rows[r] = (model, word) that model's n_wk row (hot words, before
// their pieces are added with atomics).
// Workgroups from clear_groups on sum the LL packets of the sweeps before
// (stage 1, see ll_packet_sum), which the clear does not touch.
__device__ __forceinline__ void ll_packet_sum(const LLArgs &a, int p);
__global__ __launch_bounds__(256) void phi_clear_kernel(uint32_t *fill, int64_t n_fill, const ModelDev *models,
                                                        const int2 *rows, int32_t n, int32_t clear_groups, LLArgs ll) {
    if ((int)blockIdx.x >= clear_groups) {
        ll_packet_sum(ll, blockIdx.x - clear_groups);
        return;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_fill; i += (int64_t)clear_groups * blockDim.x)
        fill[i] = 0;
    for (int r = blockIdx.x; r < n; r += clear_groups) {
        const int2 mw = rows[r];
        const ModelDev &m = models[mw.x];
        int32_t *row = m.nwk + (int64_t)mw.y * m.k;
        for (int j = threadIdx.x; j < m.k; j += blockDim.x) *glb(row + j) = 0;
    }
}

// ------------------------------------------------------------ phi rebuild
// One workgroup per packet (a run of phi slots of one model), one wave per
// slot: LDS histogram of the word's topics, whole rows written with plain
// 16-byte stores, hot-word pieces added with atomics.  n_k is accumulated
// per workgroup in LDS and flushed with one atomic per topic.
// accumulate == 2 (counts known to equal the counts of zp): rare words
// (fewer than K / 16 tokens) clear only the entries of their previous
// topics and store the bins of their new ones, with every step a loop over
// the word's tokens instead of over K; every token's zp becomes its z.
// The counts are the same either way.
// Reference: LDAUpdatePhiKernel (kernel/lda_phi_kernel.cu:76-117) and the
// chunk reduction LDAUpdatePhiReduceKernel* (:162-183).

// Histogram of z[b, e) into the wave's LDS bins h; zp = z when keep_zp.
// Each lane issues eight loads before their LDS atomics, so a long word
// waits for one memory round trip per 512 tokens instead of per 64.
__device__ __forceinline__ void count_range(const uint16_t *z, uint16_t *zp, int64_t b, int64_t e, int lane,
                                            uint32_t *h, bool keep_zp) {
    for (int64_t i = b + lane; i < e; i += 8 * kWave) {
        uint16_t t[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t x = i + j * kWave;
            t[j] = x < e ? (uint16_t)load_global_u16(z + x) : 0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t x = i + j * kWave;
            if (x < e) {
                atomicAdd(&h[t[j]], 1u);
                if (keep_zp) store_global_u16(zp + x, t[j]);
            }
        }
    }
}

// Histogram into the wave's LDS bins h of a word's tokens in units j < nu
// <= 64, lane j holding unit j's z and zp pointers and range [wb, we): the
// 64-token blocks of all those units in unit order, kGather blocks' loads
// in flight before their LDS atomics; zp = z when kKeepZp (fewer blocks in
// flight then: their zp addresses take registers too).  A word of up to
// kGather blocks waits for one memory round trip whatever the number of
// its units.  The cursor (unit j, its next block's address, the tokens left
// in it) is wave-uniform, so a block costs a few scalar operations and one
// vector load at a scalar base plus the lane's byte offset.
template <bool kKeepZp, int kGather>
__device__ __forceinline__ void count_blocks(const uint16_t *zl, uint16_t *zpl, int nu, int64_t wb, int64_t we,
                                             int lane, uint32_t *h) {
    constexpr uint32_t kNone = 0xffffffffu;  // no token in this lane of the block
    int j = 0;
    const char *zb = nullptr;
    char *zpb = nullptr;
    int64_t left = 0;
    auto seek = [&]() {  // first unit from j on with tokens left
        for (; j < nu; ++j) {
            const int64_t p = readlane64(wb, j);
            left = readlane64(we, j) - p;
            if (left > 0) {
                zb = reinterpret_cast<const char *>(readlane64((int64_t)zl, j)) + 2 * p;
                if (kKeepZp) zpb = reinterpret_cast<char *>(readlane64((int64_t)zpl, j)) + 2 * p;
                return;
            }
        }
    };
    seek();
    const uint32_t off = 2u * (uint32_t)lane;
    while (j < nu) {
        uint32_t t[kGather];
        char *zq[kGather];
#pragma unroll
        for (int q = 0; q < kGather; ++q) {
            t[q] = kNone;
            zq[q] = nullptr;
            if (j < nu) {
                if ((int64_t)lane < left) t[q] = load_global_u16_at(zb, off);
                if (kKeepZp) zq[q] = zpb;
                zb += 2 * kWave;
                zpb += 2 * kWave;
                left -= kWave;
                if (left <= 0) {
                    ++j;
                    seek();
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kGather; ++q)
            if (t[q] != kNone) {
                atomicAdd(&h[t[q]], 1u);
                if (kKeepZp) store_global_u16_at(zq[q], off, (uint16_t)t[q]);
            }
    }
}

// Body of the phi rebuild for workgroup packet `packet`, on the caller's LDS.
// A wave loads its next item's slot descriptor and word ranges while it
// counts the current one (the descriptor two items ahead); the ranges of a
// slot's first 64 units come through per-lane word_idx pointers, loaded
// again only when the slot's unit range changes (every whole-word item of a
// packet has the same one).
__device__ __forceinline__ void phi_rebuild_body(const PhiArgs &a, int packet, uint32_t (*hist_s)[kMaxK],
                                                 uint32_t *nk_s) {
    const Packet pk = a.packets[packet];
    const ModelDev &m = a.models[pk.owner];
    const int K = m.k;
    const bool keep_zp = a.accumulate == 2;
    for (int k = threadIdx.x; k < K; k += blockDim.x) nk_s[k] = 0;
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    uint32_t *h = hist_s[wave];
    int ulo = -1, uhi = -1;
    // lane j: word_idx, z and zp of unit ulo + j
    const int64_t *wil = nullptr;
    const uint16_t *zl = nullptr;
    uint16_t *zpl = nullptr;
    // lane j < min(64, unit count): the slot word's range in unit unit_lo + j
    auto load_ranges = [&](const PhiSlot &s, int64_t &b, int64_t &e) {
        b = e = 0;
        if (s.split) return;
        if (s.unit_lo != ulo || s.unit_hi != uhi) {
            ulo = s.unit_lo;
            uhi = s.unit_hi;
            if (lane < uhi - ulo) {
                const CULDA_AS1 UnitDev &u = *glb(a.units + ulo + lane);
                wil = u.word_idx;
                zl = u.z;
                zpl = u.zp;
            }
        }
        if (lane < uhi - ulo) {
            b = *glb(wil + s.word);
            e = *glb(wil + s.word + 1);
        }
    };
    constexpr int step = kWavesPerGroup;
    int si = pk.first + wave;
    PhiSlot s1{}, s2{};
    int64_t b1 = 0, e1 = 0;
    if (si < pk.last) {
        s1 = a.slots[si];
        load_ranges(s1, b1, e1);
    }
    if (si + step < pk.last) s2 = a.slots[si + step];
    for (; si < pk.last; si += step) {
        const PhiSlot sl = s1;
        const int64_t wb = b1, we = e1;
        if (si + step < pk.last) {
            s1 = s2;
            load_ranges(s1, b1, e1);
            if (si + 2 * step < pk.last) s2 = a.slots[si + 2 * step];
        }
        int32_t *row = m.nwk + (int64_t)sl.word * K;
        // fn(unit, b, e) over the word's token range [b, e) in every unit of
        // the slot, in unit order (ranges of units past the first 64 loaded
        // here, one unit per lane)
        auto each_unit = [&](auto &&fn) {
            for (int g0 = sl.unit_lo; g0 < sl.unit_hi; g0 += kWave) {
                const int cnt = sl.unit_hi - g0 < kWave ? sl.unit_hi - g0 : kWave;
                int64_t b = wb, e = we;
                if (g0 != sl.unit_lo) {
                    b = e = 0;
                    if (lane < cnt) {
                        const int64_t *wi = glb(a.units + g0 + lane)->word_idx;
                        b = *glb(wi + sl.word);
                        e = *glb(wi + sl.word + 1);
                    }
                }
                for (int j = 0; j < cnt; ++j) fn(g0 + j, readlane64(b, j), readlane64(e, j));
            }
        };
        const int nu = sl.unit_hi - sl.unit_lo;
        if (!sl.split) {
            int64_t n = 0;
            each_unit([&](int, int64_t b, int64_t e) { n += e - b; });
            if (n == 0 && a.accumulate != 0) continue;  // nothing to add, row already right
            if (keep_zp && n * 16 < K) {
                // Every loop runs over the word's tokens, not over K.  The
                // row holds the counts of zp: entries of old topics that
                // have no new token are cleared, entries of new topics get
                // their new counts -- disjoint stores, so no ordering
                // between them.
                each_unit([&](int u, int64_t b, int64_t e) {
                    const uint16_t *z = a.units[u].z, *zp = a.units[u].zp;
                    for (int64_t i = b + lane; i < e; i += kWave) {
                        h[load_global_u16(z + i)] = 0;
                        h[load_global_u16(zp + i)] = 0;
                    }
                });
                wave_sync();
                each_unit([&](int u, int64_t b, int64_t e) {
                    const uint16_t *z = a.units[u].z;
                    for (int64_t i = b + lane; i < e; i += kWave) {
                        const uint32_t t = load_global_u16(z + i);
                        atomicAdd(&h[t], 1u);
                        atomicAdd(&nk_s[t], 1u);
                    }
                });
                wave_sync();
                each_unit([&](int u, int64_t b, int64_t e) {
                    const uint16_t *z = a.units[u].z;
                    uint16_t *zp = a.units[u].zp;
                    for (int64_t i = b + lane; i < e; i += kWave) {
                        const uint16_t t = (uint16_t)load_global_u16(z + i), p = (uint16_t)load_global_u16(zp + i);
                        if (h[p] == 0) *glb(row + p) = 0;
                        *glb(row + t) = (int)h[t];
                        store_global_u16(zp + i, t);
                    }
                });
                wave_sync();
                continue;
            }
        }
        for (int k = lane * 4; k < K; k += 4 * kWave) *reinterpret_cast<uint4 *>(h + k) = make_uint4(0u, 0u, 0u, 0u);
        wave_sync();
        const uint16_t *cz = zl;
        uint16_t *czp = zpl;
        int cn = nu;
        int64_t cb = wb, ce = we;
        if (sl.split) {
            const CULDA_AS1 UnitDev &u = *glb(a.units + sl.unit_lo);
            cz = u.z;
            czp = u.zp;
            cn = 1;
            cb = sl.begin;
            ce = sl.end;
        }
        if (cn <= kWave) {
            if (keep_zp)
                count_blocks<true, 8>(cz, czp, cn, cb, ce, lane, h);
            else
                count_blocks<false, 16>(cz, czp, cn, cb, ce, lane, h);
        } else {
            each_unit([&](int u, int64_t b, int64_t e) { count_range(a.units[u].z, a.units[u].zp, b, e, lane, h, keep_zp); });
        }
        wave_sync();
        if (!sl.split) {
            // whole row with 16-byte stores; n_k from the same bins
            for (int k = lane * 4; k < K; k += 4 * kWave) {
                const uint4 v = *reinterpret_cast<const uint4 *>(h + k);
                int4 o = make_int4((int)v.x, (int)v.y, (int)v.z, (int)v.w);
                CULDA_AS1 int4 *dst = reinterpret_cast<CULDA_AS1 int4 *>(glb(row + k));
                if (a.accumulate == 1) {
                    const int4 old = *dst;
                    o.x += old.x;
                    o.y += old.y;
                    o.z += old.z;
                    o.w += old.w;
                }
                *dst = o;
                if (v.x) atomicAdd(&nk_s[k], v.x);
                if (v.y) atomicAdd(&nk_s[k + 1], v.y);
                if (v.z) atomicAdd(&nk_s[k + 2], v.z);
                if (v.w) atomicAdd(&nk_s[k + 3], v.w);
            }
        } else {
            for (int k = lane; k < K; k += kWave) {
                const uint32_t v = h[k];
                if (v) {
                    atomic_add_global(row + k, (int)v);
                    atomicAdd(&nk_s[k], v);
                }
            }
        }
        wave_sync();
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const uint32_t v = nk_s[k];
        if (v) atomic_add_global(m.nk + k, (int)v);
    }
}

__global__ __launch_bounds__(256) void phi_rebuild_kernel(PhiArgs a) {
    __shared__ __attribute__((aligned(16))) uint32_t hist_s[kWavesPerGroup][kMaxK];
    __shared__ uint32_t nk_s[kMaxK];
    phi_rebuild_body(a, blockIdx.x, hist_s, nk_s);
}

// ------------------------------------------------------------ rare tiles
// Rare words of one-unit models: their n_wk rows are brought up to date
// only when something reads them.  After a sweep (kMaterialize false) one
// wave per tile of up to 64 tokens adds each token's topic to n_k (LDS,
// flushed per packet) and sets zp = z -- the sampler stages such a word's
// phi from the zp topics of its tokens, all in its one slot, without the
// row.  Before a read of n_wk (kMaterialize true) the rows move from the
// counts of zm to the counts of zp: a token's word is its start_mask rank,
// its new bin the number of lanes with the same (word, zp), and a zm topic
// that no lane keeps is cleared -- the cleared and the stored entries are
// disjoint, so no ordering between the stores; zm = zp.
template <bool kMaterialize>
__device__ __forceinline__ void phi_rare_body(const RareArgs &a, int packet, uint32_t *nk_s) {
    const Packet pk = a.packets[packet];
    const ModelDev &m = a.models[pk.owner];
    const int K = m.k;
    if (!kMaterialize) {
        for (int k = threadIdx.x; k < K; k += blockDim.x) nk_s[k] = 0;
        __syncthreads();
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int ti = pk.first + wave; ti < pk.last; ti += kWavesPerGroup) {
        const RareTile t = a.tiles[ti];
        const UnitDev &u = a.units[t.unit];
        const bool live = lane < t.count;
        if (!kMaterialize) {
            if (live) {
                const uint16_t zt = (uint16_t)load_global_u16(u.z + t.begin + lane);
                store_global_u16(u.zp + t.begin + lane, zt);
                atomicAdd(&nk_s[zt], 1u);
            }
            continue;
        }
        uint32_t zt = 0, zo = 0;
        if (live) {
            zt = load_global_u16(u.zp + t.begin + lane);
            zo = load_global_u16(u.zm + t.begin + lane);
        }
        // word of present-word rank q: lane r (word w0 + r present) pushes
        // w0 + r to lane rank(r); absent words push to lane 63, which no
        // rank reads unless all 64 words are present
        const bool here = (t.word_mask >> lane) & 1ull;
        const int dst = here ? __popcll(t.word_mask & ((1ull << lane) - 1ull)) : 63;
        const int by_rank = __builtin_amdgcn_ds_permute(dst << 2, t.w0 + lane);
        const uint32_t rank = (uint32_t)__popcll(t.start_mask & ((2ull << lane) - 1ull)) - 1u;
        const int word = __builtin_amdgcn_ds_bpermute((int)(rank << 2), by_rank);
        const uint32_t key = (rank << 16) | zt, okey = (rank << 16) | zo;
        uint32_t c = 0;
        bool kept = false;
        for (int j = 0; j < t.count; ++j) {
            const uint32_t kj = __builtin_amdgcn_readlane(key, j);
            c += kj == key;
            kept |= kj == okey;
        }
        if (live) {
            int32_t *row = m.nwk + (int64_t)word * K;
            if (!kept) *glb(row + zo) = 0;
            *glb(row + zt) = (int32_t)c;
            store_global_u16(u.zm + t.begin + lane, (uint16_t)zt);
        }
    }
    if (!kMaterialize) {
        __syncthreads();
        for (int k = threadIdx.x; k < K; k += blockDim.x) {
            const uint32_t v = nk_s[k];
            if (v) atomic_add_global(m.nk + k, (int)v);
        }
    }
}

template <bool kMaterialize>
__global__ __launch_bounds__(256) void phi_rare_kernel(RareArgs a) {
    __shared__ uint32_t nk_s[kMaxK];
    phi_rare_body<kMaterialize>(a, blockIdx.x, nk_s);
}

// The count update after a sweep of one-unit models in one launch: the first
// p.num_packets workgroups rebuild word slots, the rest take the rare tiles'
// n_k / zp pass.  The two touch disjoint words and zp ranges, and their n_k
// atomics commute, so the counts equal those of the two launches in turn.
__global__ __launch_bounds__(256) void phi_update_kernel(PhiArgs p, RareArgs r) {
    __shared__ __attribute__((aligned(16))) uint32_t hist_s[kWavesPerGroup][kMaxK];
    __shared__ uint32_t nk_s[kMaxK];
    if ((int)blockIdx.x < p.num_packets)
        phi_rebuild_body(p, blockIdx.x, hist_s, nk_s);
    else
        phi_rare_body<false>(r, (int)blockIdx.x - p.num_packets, nk_s);
}

// den_k = fl(fl(n_k) + fl(beta * fl(V))) for every model of the table
// (kernel/lda_phi_kernel.cu:26); one workgroup per model.
// Head denominators; with ll.num_packets > 0 also the LL stage 2 of every
// model (the packet sums fetched by all threads at once, then added in
// packet order by one), and with advance (one workgroup) the sweep counter
// moves to the next sweep after the LL of this one is stored.
__global__ __launch_bounds__(256) void den_kernel(const ModelDev *models, int32_t num_models, LLArgs ll,
                                                  uint32_t *advance) {
    __shared__ double ps[256], pn[256];
    if (ll.num_packets > 0) {
        const uint32_t sweep = *ll.sweep;
        for (int mi = blockIdx.x; mi < ll.num_models; mi += gridDim.x) {
            const int p0 = ll.model_packets[2 * mi], p1 = ll.model_packets[2 * mi + 1];
            double acc = 0.0, nz = 0.0;
            for (int c = p0; c < p1; c += 256) {
                __syncthreads();
                if (c + (int)threadIdx.x < p1) {
                    ps[threadIdx.x] = ll.scratch[c + threadIdx.x];
                    pn[threadIdx.x] = ll.nnz_scratch[c + threadIdx.x];
                }
                __syncthreads();
                if (threadIdx.x == 0)
                    for (int j = 0; j < 256 && c + j < p1; ++j) {
                        acc += ps[j];
                        nz += pn[j];
                    }
            }
            if (threadIdx.x == 0) {
                const size_t h = (size_t)(sweep % kLLHistory) * ll.num_models + mi;
                ll.history[h] = acc;
                ll.nnz_history[h] = nz;
            }
        }
    }
    if (advance && threadIdx.x == 0) *advance = *advance + 1u;  // one workgroup: after its LL store
    for (int mi = blockIdx.x; mi < num_models; mi += gridDim.x) {
        const ModelDev &m = models[mi];
        const float bv = m.beta * (float)m.num_words;
        for (int i = threadIdx.x; i < m.k; i += blockDim.x) {
            const float d = (float)*glb(m.nk + i) + bv;
            *glb(m.den + i) = d;
            *glb(m.zero_phi + i) = m.beta / d;  // = fl(fl(0 + beta) / d)
            *glb(m.rden + i) = 1.0 / (double)d;
        }
    }
}

// Materialised head, for downloads and save hooks only (the sampler forms
// it on the fly while staging a word's row).
__global__ void phi_head_kernel(const int32_t *nwk, const float *den, float *head, int64_t n, int32_t k, float beta) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        head[i] = ((float)nwk[i] + beta) / den[i % k];
}

// ------------------------------------------------------------ theta rebuild
// A doc's theta entries from its LDS counts and the bitmap of its topics
// (bm: nw words, cleared here): the nonzero topics in ascending order, 64
// per step, lane r of a step taking the topic of the r-th set bit (a binary
// search over the words' popcount prefixes, then over the word's halves);
// a count above kThetaMaxCount takes several entries (kThetaMaxCount each,
// the remainder last), placed by a wave prefix sum.  take(k) returns topic
// k's count and clears it, put(p, e) stores entry e at position p.
// Returns the number of entries.
template <class Take, class Put>
__device__ __forceinline__ uint32_t theta_compact(uint32_t *bm, uint32_t nw, int lane, Take take, Put put) {
    // lane j < nw: bitmap word j and the number of set bits before it
    // (lanes past the last word: the total, which no rank reaches)
    const uint32_t w = (uint32_t)lane < nw ? bm[lane] : 0u;
    if ((uint32_t)lane < nw) bm[lane] = 0u;
    const uint32_t pc = (uint32_t)__popc(w);
    const uint32_t incl_w = wave_add_scan_u32(pc);
    const uint32_t ex = incl_w - pc;
    const uint32_t nz = (uint32_t)__builtin_amdgcn_readlane((int)incl_w, kWave - 1);
    uint32_t base = 0;
    for (uint32_t r0 = 0; r0 < nz; r0 += kWave) {
        // every lane searches (the permutes read from all lanes); only
        // ranks below nz are used
        const uint32_t r = r0 + (uint32_t)lane;
        uint32_t j = 0;  // the last word whose prefix is at most r
#pragma unroll
        for (uint32_t step = 16; step >= 1; step >>= 1) {
            const uint32_t ej = (uint32_t)__builtin_amdgcn_ds_bpermute((int)((j + step) << 2), (int)ex);
            if (ej <= r) j += step;
        }
        const uint32_t wj = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(j << 2), (int)w);
        uint32_t q = r - (uint32_t)__builtin_amdgcn_ds_bpermute((int)(j << 2), (int)ex);
        // bit index of the q-th set bit of wj
        uint32_t x = wj, bit = 0;
#pragma unroll
        for (uint32_t width = 16; width >= 1; width >>= 1) {
            const uint32_t low = (uint32_t)__popc(x & ((1u << width) - 1u));
            if (q >= low) {
                q -= low;
                x >>= width;
                bit += width;
            }
        }
        const uint32_t k = j * 32u + bit;
        const uint32_t v = r < nz ? take(k) : 0u;
        if (__ballot(v > kThetaMaxCount) == 0ull) {
            if (r < nz) put(base + (uint32_t)lane, (uint16_t)((v << kThetaCountShift) | k));
            base += min(nz - r0, (uint32_t)kWave);
        } else {
            const uint32_t ne = (v + kThetaMaxCount - 1u) / kThetaMaxCount;  // 0 past nz
            const uint32_t incl = wave_add_scan_u32(ne);
            const uint32_t pos = base + incl - ne;
            for (uint32_t c0 = 0; c0 < ne; ++c0) {
                const uint32_t c = c0 + 1u < ne ? kThetaMaxCount : v - kThetaMaxCount * (ne - 1u);
                put(pos + c0, (uint16_t)((c << kThetaCountShift) | k));
            }
            base += (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
        }
    }
    return base;
}

// One workgroup per packet (a run of docs of one unit), one wave per doc:
// LDS histogram over the doc's tokens (gathered through rev_idx) and a
// bitmap of the topics they touch, then theta_compact: a doc costs
// ceil(nonzero topics / 64) steps instead of K / 64.  Counters and bitmap
// words are cleared as they are read.
// Reference: LDAUpdateThetaKernel (kernel/lda_theta_kernel.cu:5-87).
__global__ __launch_bounds__(256) void theta_rebuild_kernel(ThetaArgs a) {
    __shared__ __attribute__((aligned(16))) uint32_t hist_s[kWavesPerGroup][kMaxK];
    __shared__ uint32_t bits_s[kWavesPerGroup][kMaxK / 32];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t *const h = hist_s[wave];
    uint32_t *const bm = bits_s[wave];
    for (int k = lane; k < kMaxK; k += kWave) h[k] = 0u;
    if (lane < kMaxK / 32) bm[lane] = 0u;
    wave_sync();
    // a grid smaller than the packet list takes the packets in turn (a
    // rebuild beside a sweep holds a bounded share of the CUs)
    for (int pi = blockIdx.x; pi < a.num_packets; pi += gridDim.x) {
    const Packet pk = a.packets[pi];
    const UnitDev &un = a.units[pk.owner];
    const ModelDev &m = a.models[un.model];
    const int K = m.k;
    const uint32_t nw = (uint32_t)(K + 31) / 32u;  // bitmap words, at most 32
    for (int d = pk.first + wave; d < pk.last; d += kWavesPerGroup) {
        const uint32_t b = un.rev_off[d], e = un.rev_off[d + 1];
        // four index loads, then four z gathers, in flight per lane
        for (uint32_t i0 = b + lane; i0 < e; i0 += 4 * kWave) {
            uint32_t t[4];
            uint16_t zz[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) t[q] = i0 + q * kWave < e ? un.rev_idx[i0 + q * kWave] : 0xffffffffu;
#pragma unroll
            for (int q = 0; q < 4; ++q) zz[q] = t[q] != 0xffffffffu ? un.z[t[q]] : (uint16_t)0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (t[q] != 0xffffffffu) {
                    atomicAdd(&h[zz[q]], 1u);
                    atomicOr(&bm[zz[q] >> 5], 1u << (zz[q] & 31u));
                }
        }
        wave_sync();
        const uint32_t start = un.meta[d].start;
        const uint32_t base = theta_compact(
            bm, nw, lane,
            [&](uint32_t k) {
                const uint32_t v = h[k];
                h[k] = 0u;
                return v;
            },
            [&](uint32_t p, uint16_t en) { un.theta[start + p] = en; });
        // zero the slab tail up to the next multiple of 16 (read by the sampler)
        if ((uint32_t)lane < ((16u - (base & 15u)) & 15u)) un.theta[start + base + lane] = 0u;
        if (lane == 0) un.meta[d].nnz = base;
        wave_sync();
    }
    }
}

// The rebuild beside a sweep, sized to fit in what the sampler leaves of a
// CU, so that it takes no wave slot a sweep could use: one wave per
// workgroup, at most 32 VGPRs (six sampler waves of 80 hold 480 of a SIMD's
// 512) and 2 KB of LDS, the doc's counts as 16-bit halves of 32-bit words --
// every doc of the unit must then be shorter than 65,536 tokens, and the
// unit below 2^30 tokens for 32-bit byte offsets (the host checks) -- plus
// the topic bitmap of theta_compact.  The counters are cleared once; each
// doc's compaction clears the ones it reads.  A packet's docs are the
// wave's, in turn.  Entries, their order and nnz are theta_rebuild_kernel's.
__global__ __launch_bounds__(64) void theta_rebuild_slim_kernel(ThetaArgs a) {
    __shared__ __attribute__((aligned(16))) uint32_t h[kMaxK / 2];
    __shared__ uint32_t bm[kMaxK / 32];
    uint16_t *const h16 = reinterpret_cast<uint16_t *>(h);
    const int lane = threadIdx.x;
    for (int i = lane; i < kMaxK / 2; i += kWave) h[i] = 0u;
    if (lane < kMaxK / 32) bm[lane] = 0u;
    for (int pi = blockIdx.x; pi < a.num_packets; pi += gridDim.x) {
        // descriptors in scalar registers, read once per packet
        const Packet pk = uniform_copy(a.packets + pi);
        const UnitDev un = uniform_copy(a.units + pk.owner);
        const int K = __builtin_amdgcn_readfirstlane(a.models[un.model].k);
        for (int d = pk.first; d < pk.last; ++d) {
            wave_sync();
            const uint32_t b = __builtin_amdgcn_readfirstlane(load_global_at(un.rev_off, 4u * d));
            const uint32_t e = __builtin_amdgcn_readfirstlane(load_global_at(un.rev_off, 4u * d + 4u));
            // four index loads, then four z gathers, in flight per lane
            for (uint32_t i0 = b + lane; i0 < e; i0 += 4 * kWave) {
                uint32_t t[4], zz[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    t[q] = i0 + q * kWave < e ? load_global_at(un.rev_idx, 4u * (i0 + q * kWave)) : 0xffffffffu;
#pragma unroll
                for (int q = 0; q < 4; ++q) zz[q] = t[q] != 0xffffffffu ? load_global_u16_at(un.z, 2u * t[q]) : 0u;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (t[q] != 0xffffffffu) {
                        atomicAdd(&h[zz[q] >> 1], 1u << ((zz[q] & 1u) * 16u));
                        atomicOr(&bm[zz[q] >> 5], 1u << (zz[q] & 31u));
                    }
            }
            wave_sync();
            const uint32_t start = __builtin_amdgcn_readfirstlane(glb(un.meta + d)->start);
            const uint32_t base = theta_compact(
                bm, (uint32_t)(K + 31) / 32u, lane,
                [&](uint32_t k) {
                    const uint32_t v = h16[k];
                    h16[k] = 0;
                    return v;
                },
                [&](uint32_t p, uint16_t en) { store_global_u16_at(un.theta, 2u * (start + p), en); });
            // zero the slab tail up to the next multiple of 16 (read by the sampler)
            if ((uint32_t)lane < ((16u - (base & 15u)) & 15u)) store_global_u16_at(un.theta, 2u * (start + base + lane), 0u);
            if (lane == 0) glb(un.meta + d)->nnz = base;
        }
    }
}

// The same rebuild for a launch that runs alone (a single-chunk model,
// whose next sweep waits for it): the histogram is cleared once per
// workgroup and each doc's compaction (theta_compact) clears what it reads.
// bits: kWavesPerGroup topic bitmaps of kMaxK / 32 words.
__device__ __forceinline__ void theta_alone_body(const ThetaArgs &a, int packet, uint32_t (*hist_s)[kMaxK],
                                                 uint32_t *bits) {
    const Packet pk = a.packets[packet];
    const UnitDev &un = a.units[pk.owner];
    const ModelDev &m = a.models[un.model];
    const int K = m.k;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t *h = hist_s[wave];
    uint32_t *const bm = bits + wave * (kMaxK / 32);
    for (int k = 4 * lane; k < K; k += 4 * kWave) *reinterpret_cast<uint4 *>(h + k) = make_uint4(0, 0, 0, 0);
    if (lane < kMaxK / 32) bm[lane] = 0u;
    for (int d = pk.first + wave; d < pk.last; d += kWavesPerGroup) {
        wave_sync();
        const uint32_t b = load_global(un.rev_off + d), e = load_global(un.rev_off + d + 1);
        const uint32_t start = glb(un.meta + d)->start;  // issued before the histogram's waits
        // four index loads, then four z gathers, in flight per lane
        for (uint32_t i0 = b + lane; i0 < e; i0 += 4 * kWave) {
            uint32_t t[4];
            uint16_t zz[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) t[q] = i0 + q * kWave < e ? load_global(un.rev_idx + i0 + q * kWave) : 0xffffffffu;
#pragma unroll
            for (int q = 0; q < 4; ++q) zz[q] = t[q] != 0xffffffffu ? (uint16_t)load_global_u16(un.z + t[q]) : (uint16_t)0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (t[q] != 0xffffffffu) {
                    atomicAdd(&h[zz[q]], 1u);
                    atomicOr(&bm[zz[q] >> 5], 1u << (zz[q] & 31u));
                }
        }
        wave_sync();
        const uint32_t base = theta_compact(
            bm, (uint32_t)(K + 31) / 32u, lane,
            [&](uint32_t k) {
                const uint32_t v = h[k];
                h[k] = 0u;
                return v;
            },
            [&](uint32_t p, uint16_t en) { store_global_u16(un.theta + start + p, en); });
        // zero the slab tail up to the next multiple of 16 (read by the sampler)
        if ((uint32_t)lane < ((16u - (base & 15u)) & 15u)) store_global_u16(un.theta + start + base + lane, 0u);
        if (lane == 0) glb(un.meta + d)->nnz = base;
        wave_sync();
    }
}

__global__ __launch_bounds__(256) void theta_rebuild_alone_kernel(ThetaArgs a) {
    __shared__ __attribute__((aligned(16))) uint32_t hist_s[kWavesPerGroup][kMaxK];
    __shared__ uint32_t bits_s[kWavesPerGroup * (kMaxK / 32)];
    theta_alone_body(a, blockIdx.x, hist_s, bits_s);
}

// A one-unit model's count update and theta rebuild in one launch (both
// read only the new z): workgroups [0, t.num_packets) rebuild theta, then
// the phi slot packets, then the rare tiles (r.mode 0).  The next sweep
// depends on both, so neither needs a stream of its own.
__global__ __launch_bounds__(256) void count_update_kernel(PhiArgs p, RareArgs r, ThetaArgs t) {
    __shared__ __attribute__((aligned(16))) uint32_t hist_s[kWavesPerGroup][kMaxK];
    __shared__ uint32_t nk_s[kMaxK];
    const int b = blockIdx.x;
    if (b < t.num_packets)
        theta_alone_body(t, b, hist_s, nk_s);  // nk_s: the theta workgroups' topic bitmaps
    else if (b < t.num_packets + p.num_packets)
        phi_rebuild_body(p, b - t.num_packets, hist_s, nk_s);
    else
        phi_rare_body<false>(r, b - t.num_packets - p.num_packets, nk_s);
}

// ------------------------------------------------------------ sampler
// One wave per slot (one word's token range of one unit).  The wave stages
// the word's head row phi_k in LDS (computed from n_wk), builds the
// Q-bucket block prefix, then takes its tokens 64 at a time: lane l loads
// the doc id, doc meta and Philox uniform of token l; the batch is sorted
// by nnz so that the eight 8-lane rows, which process eight tokens side by
// side, run S loops of similar length.  Theta groups are prefetched a
// group ahead; scans are DPP row operations.  z and the FP64 LL term are
// written per token.  The processing order never changes a result: every
// token depends only on the frozen counts and its own Philox draw.
// Persistent mode (queue != null): each workgroup claims kQueueClaim
// consecutive list positions per atomic on a shared counter until the list
// is exhausted (the host interleaves the order so that every claim holds
// slots of all lengths); otherwise workgroup b takes position b.
// Reference: LDAKernelTrain (kernel/lda_train_kernel.cu:20-271).
// kWaves waves per SIMD bound the registers (7: at most 72 VGPRs, 6: 80):
// the per-token chains of dependent loads need the occupancy more than the
// registers, unless the S-prefix rows in LDS already hold it below seven.
constexpr int kStageSteps = kMaxK / (4 * kWave);  // 16-byte loads per lane that stage a row

template <bool kPersistent, int kWaves>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kWaves))) void sample_kernel(SampleArgs a) {
    // one wave per workgroup: every LDS array sits at a fixed address, so
    // the phi lookup is the entry's topic times 4 plus a constant offset
    __shared__ __attribute__((aligned(16))) float ph[kMaxK];
    __shared__ __attribute__((aligned(16))) float rq[kWave];
    // eight rows of a.rs_stride floats: S block prefixes of each row's doc
    extern __shared__ __attribute__((aligned(16))) float rs_s[];
    const int lane = threadIdx.x;
    const uint32_t sweep = *a.sweep;
    const int rrow = lane >> 3, lr = lane & 7;
    // rr[b] = inclusive S prefix after block b of this row's doc (16-byte
    // aligned: two groups' four prefixes are one store)
    float *const rr = rs_s + rrow * a.rs_stride;
    // sixteen 4-lane rows of a.rs_stride4 floats over the same LDS, for docs
    // of at most qmax blocks
    const int qrow = lane >> 2, ql = lane & 3;
    float *const rq4 = rs_s + qrow * a.rs_stride4;
    const uint32_t qmax = (uint32_t)min(a.rs_stride4 - 1, 62);
    int pos = blockIdx.x;
    uint32_t next = 0, claim_end = 0;  // persistent: claimed positions [next, claim_end)
    while (true) {
    if (kPersistent) {
        if (next >= claim_end) {
            uint32_t q = 0;
            if (lane == 0) q = atomicAdd(a.queue, (uint32_t)kQueueClaim);
            next = __builtin_amdgcn_readfirstlane(q);  // wave-uniform: keeps descriptors in SGPRs
            claim_end = next + kQueueClaim;
        }
        pos = (int)next++;
    }
    if (pos >= a.num_slots) break;
    const int slot = a.order ? (int)__builtin_amdgcn_readfirstlane(a.order[pos]) : pos;
    const SlotDesc s = uniform_copy(a.slots + slot);
    const UnitDev un = uniform_copy(a.units + s.unit);
    const ModelDev md = uniform_copy(a.models + un.model);
    const int K = md.k;
    const float alpha = md.alpha, beta = md.beta;
    const uint16_t *theta = un.theta;
    // phi and count of the low and high 16-bit theta entries of a dword d
#define PHI_LO(d) (*reinterpret_cast<const float *>(reinterpret_cast<const char *>(ph) + (((d) << 2) & 0xffcu)))
#define PHI_HI(d) (*reinterpret_cast<const float *>(reinterpret_cast<const char *>(ph) + (((d) >> 14) & 0xffcu)))
#define COUNT_LO(d) ((float)(((d) >> kThetaCountShift) & kThetaMaxCount))
#define COUNT_HI(d) ((float)((d) >> (16 + kThetaCountShift)))

    // stage phi_k = fl(fl(fl(n_wk) + beta) / den_k), formed as the FP32
    // rounding of an FP64 product with 1 / den_k (see phi_of)
    const int32_t *row = md.nwk + (int64_t)s.word * K;
    if (a.use_zp && s.rare) {
        // A rare word's counts are those of its tokens' zp topics, and
        // every token of the word is in this slot (fewer than K / 16 <= 64
        // of them, one per lane): the zero-count row (fl(beta / den_k), the
        // same value) plus, at each zp topic, its count among the lanes --
        // the row itself need not be current.
        const int n = (int)(s.end - s.begin);
        const uint32_t t = lane < n ? load_global_u16(un.zp + s.begin + lane) : 0xffffffffu;
        // each lane's count of its topic, by LDS atomics on the row's own
        // words at the zp topics (cleared first), read back before the
        // zero-count row is staged over them
        uint32_t *const hc = reinterpret_cast<uint32_t *>(ph);
        if (lane < n) hc[t] = 0u;
        wave_sync();
        if (lane < n) atomicAdd(hc + t, 1u);
        wave_sync();
        const uint32_t c = lane < n ? hc[t] : 0u;
        wave_sync();
        for (int k = lane * 4; k < K; k += 4 * kWave)
            *reinterpret_cast<u32x4 *>(ph + k) = load_global4(md.zero_phi + k);
        wave_sync();
        if (lane < n) ph[t] = phi_of((float)c + beta, *glb(md.rden + t));
    } else {
        // the loads of kBatch 256-topic steps of the row and of den issued
        // before their first use (all of them: one memory round trip per
        // row; the persistent loop keeps more registers live, so one)
        constexpr int kBatch = kPersistent ? 1 : kStageSteps;
#pragma unroll
        for (int i0 = 0; i0 < kStageSteps; i0 += kBatch) {
            if (lane * 4 + i0 * 4 * kWave >= K) break;
            u32x4 cr[kBatch];
            f64x2 ra[kBatch], rb[kBatch];
#pragma unroll
            for (int i = 0; i < kBatch; ++i) {
                const int k = lane * 4 + (i0 + i) * 4 * kWave;
                if (k < K) {
                    cr[i] = load_global4(row + k);
                    ra[i] = *glb(reinterpret_cast<const f64x2 *>(md.rden + k));
                    rb[i] = *glb(reinterpret_cast<const f64x2 *>(md.rden + k + 2));
                }
            }
#pragma unroll
            for (int i = 0; i < kBatch; ++i) {
                const int k = lane * 4 + (i0 + i) * 4 * kWave;
                if (k < K) {
                    ph[k + 0] = phi_of((float)(int32_t)cr[i].x + beta, ra[i].x);
                    ph[k + 1] = phi_of((float)(int32_t)cr[i].y + beta, ra[i].y);
                    ph[k + 2] = phi_of((float)(int32_t)cr[i].z + beta, rb[i].x);
                    ph[k + 3] = phi_of((float)(int32_t)cr[i].w + beta, rb[i].y);
                }
            }
        }
    }
    wave_sync();
    // Q-bucket block totals (balanced tree of 16), then a 64-lane
    // Kogge-Stone scan of the totals
    const int nbq = K >> 4;
    float T = 0.0f;
    if (lane < nbq) {
        const float *v = ph + lane * kBlock;
        float q[16];
#pragma unroll
        for (int i = 1; i < 16; i += 2) q[i] = alpha * v[i];
        // pair sums fma(alpha, v_2m, fl(alpha v_2m+1)), as every block total
        const float t01 = __builtin_fmaf(alpha, v[0], q[1]), t23 = __builtin_fmaf(alpha, v[2], q[3]);
        const float t45 = __builtin_fmaf(alpha, v[4], q[5]), t67 = __builtin_fmaf(alpha, v[6], q[7]);
        const float t89 = __builtin_fmaf(alpha, v[8], q[9]), tab = __builtin_fmaf(alpha, v[10], q[11]);
        const float tcd = __builtin_fmaf(alpha, v[12], q[13]), tef = __builtin_fmaf(alpha, v[14], q[15]);
        const float u03 = t01 + t23, u47 = t45 + t67, u8b = t89 + tab, ucf = tcd + tef;
        T = (u03 + u47) + (u8b + ucf);
    }
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const float y = __shfl_up(T, d, kWave);
        if (lane >= d) T = T + y;
    }
    // past the last block: -inf, which no target is below (the block
    // search reads four prefixes at a time without bounds per prefix)
    rq[lane] = lane < nbq ? T : -__builtin_inff();
    const float Q = __shfl(T, nbq - 1, kWave);
    wave_sync();

    double ll = 0.0;
    uint32_t gathered = 0;  // theta entries this lane's tokens read (the slot's total fits 32 bits)
    // Philox blocks [pc0, pc0 + 64), block pc0 + l's four words (top 24
    // bits) in lane l: token g draws from word g & 3 of block g >> 2, so one
    // evaluation per lane serves up to four batches of the slot
    int64_t pc0 = -1;
    uint32_t pw0 = 0, pw1 = 0, pw2 = 0, pw3 = 0;
    for (int64_t base = s.begin; base < s.end; base += kWave) {
        const int ntok = (int)((s.end - base) < kWave ? (s.end - base) : kWave);
        const bool valid = lane < ntok;
        const int64_t g0 = base + s.gdelta;
        if (pc0 < 0 || ((g0 + ntok - 1) >> 2) >= pc0 + kWave) {
            pc0 = g0 >> 2;
            const u64x4 x = philox4x64_10((uint64_t)(pc0 + lane), sweep, kPurposeSweep, 0, md.seed, 0);
            pw0 = uniform_bits(x.v[0]);
            pw1 = uniform_bits(x.v[1]);
            pw2 = uniform_bits(x.v[2]);
            pw3 = uniform_bits(x.v[3]);
        }
        float u;
        {
            const int64_t g = g0 + lane;
            const int src = (int)(((g >> 2) - pc0) & (kWave - 1));  // < 64 for valid lanes
            const uint32_t w0 = __shfl(pw0, src, kWave), w1 = __shfl(pw1, src, kWave);
            const uint32_t w2 = __shfl(pw2, src, kWave), w3 = __shfl(pw3, src, kWave);
            const uint32_t k = (uint32_t)g & 3u;
            u = bits24_to_uniform(k == 0 ? w0 : k == 1 ? w1 : k == 2 ? w2 : w3);
        }
        uint32_t mstart = 0, mnnz = 0, mnd = 0;
        if (valid) {
            const int64_t i = base + lane;
            const uint32_t doc = load_global(un.tokens + i);
            const u32x4 m = load_global4(un.meta + doc);
            mstart = m.x;
            mnnz = m.y;
            mnd = m.z;
        }
        // Processing order: a counting sort of the valid tokens (lanes
        // 0..ntok-1) by the number of 16-entry S blocks of their doc
        // (capped at 63), descending, so that the rows of a step run S loops
        // of similar length.  The histogram sits in the first 64 words of the
        // S-prefix rows (free between batches): lane l's count of tokens
        // with key l, then the number of tokens with a larger key.  Equal
        // keys take the order of the LDS atomics (the order never changes a
        // result).  The batch's data moves between lanes by permutes, not
        // through LDS arrays: LDS bounds the waves per CU.
        const uint32_t key = min((mnnz + 15u) >> 4, 63u);
        uint32_t *const hist = reinterpret_cast<uint32_t *>(rs_s);
        hist[lane] = 0u;
        wave_sync();
        uint32_t pe = 0;
        if (valid) pe = atomicAdd(hist + key, 1u);
        wave_sync();
        hist[lane] = (uint32_t)ntok - wave_add_scan_u32(hist[lane]);
        wave_sync();
        // valid lanes take positions 0..ntok-1, the others keep their own
        const uint32_t rank = valid ? hist[key] + pe : (uint32_t)lane;
        wave_sync();
        // lane p holds the token drawn at position p (a forward permute)
        const uint32_t pstart = (uint32_t)__builtin_amdgcn_ds_permute((int)(rank << 2), (int)mstart);
        const uint32_t pnnz = (uint32_t)__builtin_amdgcn_ds_permute((int)(rank << 2), (int)mnnz);
        const float pu = __builtin_bit_cast(float, __builtin_amdgcn_ds_permute((int)(rank << 2), __builtin_bit_cast(int, u)));
        uint32_t ztop = 0;  // lane p: topic and S + Q of position p
        float zsq = 0.0f;
        // Steps of eight 8-lane rows while position r's doc has more blocks
        // than a 4-lane row's prefix row holds (the longest docs come
        // first), then steps of sixteen 4-lane rows, which share the
        // per-step work among twice the tokens.
        int r = 0;
        for (; r < ntok; r += 8) {
            if ((((uint32_t)__builtin_amdgcn_readlane((int)pnnz, r) + 15u) >> 4) <= qmax) break;
            const int p = r + rrow;
            const uint32_t start = __shfl(pstart, p, kWave);
            const uint32_t nnz = __shfl(pnnz, p, kWave);
            const float uu = __shfl(pu, p, kWave);
            uint32_t topic = 0;
            float SQ = 0.0f;
            if (p < ntok) {
                // S = sum over the doc's theta entries, 16 per block; every
                // block lies inside the doc's 16-aligned slab, entries past
                // nnz are zero (the slab tail), so their products are exact
                // zeros.  A block total is the balanced tree of its 8 pair
                // sums (each an FMA of the even product into the odd one):
                // lane m holds entries
                // 4m .. 4m + 3 of a group of two blocks (one 8-byte load;
                // block m >> 2), forms the tree's first two levels in
                // registers, and DPP row_shr 1 and 2 finish the two totals
                // at lanes 3 and 7; lane 7 carries R_{b+1} = fl(R_b + T_b)
                // in block order (row_shr 4 brings lane 3's total).  Groups
                // (32 entries) are loaded one group ahead; the last group is
                // loaded whole: loads past the last block stay inside the
                // theta array (kThetaPad entries of tail padding), and rr
                // entries past nb may hold sums of such blocks, which
                // nothing reads.
                const uint32_t nb = (nnz + 15u) >> 4;
                const uint32_t ng = (nb + 1u) >> 1;
                // byte offset of this lane's entries in the first group (a
                // chunk's theta holds fewer than 2^31 entries)
                const uint32_t tp = (start + 4u * (uint32_t)lr) * 2u;
                auto group = [&](u32x2 d, float R0, float &r0, float &r1) {
                    const float f0 = PHI_LO(d.x), f1 = PHI_HI(d.x), f2 = PHI_LO(d.y), f3 = PHI_HI(d.y);
                    const float c0 = COUNT_LO(d.x), c1 = COUNT_HI(d.x), c2 = COUNT_LO(d.y), c3 = COUNT_HI(d.y);
                    float y = __builtin_fmaf(c0, f0, c1 * f1) + __builtin_fmaf(c2, f2, c3 * f3);
                    y = y + dpp_f<0x111>(y);
                    y = y + dpp_f<0x112>(y);
                    r0 = R0 + dpp_f<0x114>(y);
                    r1 = r0 + y;
                };
                // two groups per pass, each loaded while the one before is
                // summed (no register copy between passes)
                u32x2 ca = load_global2_at(theta, tp);
                float R = 0.0f, r0, r1;
                uint32_t g = 0, o = tp;  // o: byte offset of group g's entries
                for (; g + 2 < ng; g += 2, o += 128u) {
                    const u32x2 cb = load_global2_at_imm<64>(theta, o);
                    float r2, r3;
                    group(ca, R, r0, r1);
                    ca = load_global2_at_imm<128>(theta, o);
                    group(cb, r1, r2, r3);
                    if (lr == 7) *reinterpret_cast<float4 *>(rr + (g << 1)) = make_float4(r0, r1, r2, r3);
                    R = r3;
                }
                if (g + 1 < ng) {
                    const u32x2 cb = load_global2_at_imm<64>(theta, o);
                    group(ca, R, r0, r1);
                    if (lr == 7) *reinterpret_cast<float2 *>(rr + (g << 1)) = make_float2(r0, r1);
                    R = r1;
                    ca = cb;
                    ++g;
                }
                // last group: one or two blocks
                group(ca, R, r0, r1);
                if (lr == 7) *reinterpret_cast<float2 *>(rr + (g << 1)) = make_float2(r0, r1);
                wave_sync();
                const float S = rr[nb - 1];  // nb >= 1: a token's doc has an entry
                // t = fl(u * (S + Q)); the S bucket takes t < S with target
                // t, the Q bucket target fl(t - S)
                SQ = S + Q;
                const float tu = uu * SQ;
                const bool sb = tu < S;
                const float t = sb ? tu : tu - S;
                const float *rsel = sb ? rr : rq;  // inclusive prefix after each block
                const int nbb = sb ? (int)nb : nbq;
                // first block b with t < rsel[b]: lane m tests blocks
                // q0 + 4m .. q0 + 4m + 3, the row takes the least hit.
                // Prefixes read past the last block (at most three) need no
                // test of their own: rq holds -inf there, and in the S
                // bucket t < S = rr[nb - 1], so a hit at or before nb - 1
                // exists and any later one loses the minimum.
                int bsel = -1;
                for (int q0 = 0; q0 < nbb; q0 += 32) {
                    const int b = q0 + 4 * lr;
                    uint32_t hit = 0xffu;
                    if (b < nbb) {
                        // one 16-byte read, conditions without short-circuit
                        const float4 v = *reinterpret_cast<const float4 *>(rsel + b);
                        hit = (t < v.w) ? 4u * lr + 3u : hit;
                        hit = (t < v.z) ? 4u * lr + 2u : hit;
                        hit = (t < v.y) ? 4u * lr + 1u : hit;
                        hit = (t < v.x) ? 4u * lr : hit;
                    }
                    hit = row8_min(hit);
                    if (hit != 0xffu) {
                        bsel = q0 + (int)hit;
                        break;
                    }
                }
                if (bsel < 0) {
                    topic = sb ? (load_global_u16_at(theta, (start + nnz - 1u) * 2u) & kThetaTopicMask) : (uint32_t)(K - 1);
                } else {
                    // the block's 16 products, two per lane (entries j, j + 1);
                    // entries past nnz are zero-count padding (exact zero
                    // products) and never chosen: their prefixes, summed in
                    // another order, may round above the last entry's
                    const uint32_t j = ((uint32_t)bsel << 4) + 2u * (uint32_t)lr;
                    uint32_t e2 = 0;
                    float alo, blo, vlo, vhi;
                    if (sb) {
                        e2 = load_global_at(theta, (start + j) * 2u);
                        alo = COUNT_LO(e2);
                        blo = PHI_LO(e2);
                        vhi = COUNT_HI(e2) * PHI_HI(e2);
                    } else {
                        const float2 f = *reinterpret_cast<const float2 *>(ph + j);
                        alo = alpha;
                        blo = f.x;
                        vhi = alpha * f.y;
                    }
                    prefix16x2(alo, blo, vlo, vhi, lr);
                    // prefix before the block (0 for the first: 0 + v = v,
                    // the products and sums being nonnegative)
                    const float rb = bsel ? (sb ? rr : rq)[bsel - 1] : 0.0f;
                    const float Clo = rb + vlo, Chi = rb + vhi;
                    const uint32_t lim = sb ? nnz : 0xffffffffu;
                    uint32_t hit = ((j + 1u < lim) & (t < Chi)) ? 2u * lr + 1u : 0xffu;
                    hit = row8_min(((j < lim) & (t < Clo)) ? 2u * lr : hit);
                    int msel;
                    if (hit != 0xffu) {
                        msel = (int)hit;
                    } else {
                        const int nvalid = (int)nnz - (bsel << 4);
                        msel = sb ? (nvalid < 16 ? nvalid : 16) - 1 : 15;
                    }
                    const uint32_t ew = __shfl(e2, (rrow << 3) + (msel >> 1), kWave);
                    topic = sb ? (((msel & 1) ? (ew >> 16) : ew) & kThetaTopicMask) : (uint32_t)((bsel << 4) + msel);
                }
            }
            // position r + m's result is in the first lane of row m
            const int src = ((lane - r) & 7) << 3;
            const uint32_t zt = __shfl(topic, src, kWave);
            const float zs = __shfl(SQ, src, kWave);
            if ((unsigned)(lane - r) < 8u) {
                ztop = zt;
                zsq = zs;
            }
        }
        for (; r < ntok; r += 16) {
            const int p = r + qrow;  // >= 64 only past ntok
            const uint32_t start = __shfl(pstart, p, kWave);
            const uint32_t nnz = __shfl(pnnz, p, kWave);
            const float uu = __shfl(pu, p, kWave);
            uint32_t topic = 0;
            float SQ = 0.0f;
            if (p < ntok) {
                // S over the doc's blocks: lane q of the 4-lane row holds
                // entries 4q .. 4q + 3 of block b (one 8-byte load), forms
                // the first two tree levels, and DPP row_shr 1 and 2 finish
                // the block total at lane 3 -- the same balanced tree as
                // the 8-lane rows -- which carries R_{b+1} = fl(R_b + T_b).
                // Blocks are loaded one ahead; none past nb.
                const uint32_t nb = (nnz + 15u) >> 4;
                const uint32_t tp = (start + 4u * (uint32_t)ql) * 2u;
                auto block = [&](u32x2 d, float R0) {
                    const float f0 = PHI_LO(d.x), f1 = PHI_HI(d.x), f2 = PHI_LO(d.y), f3 = PHI_HI(d.y);
                    const float c0 = COUNT_LO(d.x), c1 = COUNT_HI(d.x), c2 = COUNT_LO(d.y), c3 = COUNT_HI(d.y);
                    float y = __builtin_fmaf(c0, f0, c1 * f1) + __builtin_fmaf(c2, f2, c3 * f3);
                    y = y + dpp_f<0x111>(y);
                    y = y + dpp_f<0x112>(y);
                    return R0 + y;
                };
                u32x2 ca = load_global2_at(theta, tp);
                float R = 0.0f;
                uint32_t b = 0, o = tp;  // o: byte offset of block b's entries
                for (; b + 2 < nb; b += 2, o += 64u) {
                    const u32x2 cb = load_global2_at_imm<32>(theta, o);
                    const float r0 = block(ca, R);
                    ca = load_global2_at_imm<64>(theta, o);
                    const float r1 = block(cb, r0);
                    if (ql == 3) *reinterpret_cast<float2 *>(rq4 + b) = make_float2(r0, r1);
                    R = r1;
                }
                if (b + 1 < nb) {
                    const u32x2 cb = load_global2_at_imm<32>(theta, o);
                    const float r0 = block(ca, R);
                    const float r1 = block(cb, r0);
                    if (ql == 3) *reinterpret_cast<float2 *>(rq4 + b) = make_float2(r0, r1);
                } else {
                    const float r0 = block(ca, R);
                    if (ql == 3) rq4[b] = r0;
                }
                wave_sync();
                const float S = rq4[nb - 1];
                SQ = S + Q;
                const float tu = uu * SQ;
                const bool sb = tu < S;
                const float t = sb ? tu : tu - S;
                const float *rsel = sb ? rq4 : rq;
                const int nbb = sb ? (int)nb : nbq;
                // first block b with t < rsel[b]: lane q tests blocks q0 + 8q
                // .. q0 + 8q + 7 (two 16-byte reads), the row takes the least
                // hit (reads past the last block: as for the 8-lane rows)
                int bsel = -1;
                for (int q0 = 0; q0 < nbb; q0 += 32) {
                    const int bb = q0 + 8 * ql;
                    uint32_t hit = 0xffu;
                    if (bb < nbb) {
                        if (bb + 4 < nbb) {
                            const float4 w = *reinterpret_cast<const float4 *>(rsel + bb + 4);
                            hit = (t < w.w) ? 8u * ql + 7u : hit;
                            hit = (t < w.z) ? 8u * ql + 6u : hit;
                            hit = (t < w.y) ? 8u * ql + 5u : hit;
                            hit = (t < w.x) ? 8u * ql + 4u : hit;
                        }
                        const float4 v = *reinterpret_cast<const float4 *>(rsel + bb);
                        hit = (t < v.w) ? 8u * ql + 3u : hit;
                        hit = (t < v.z) ? 8u * ql + 2u : hit;
                        hit = (t < v.y) ? 8u * ql + 1u : hit;
                        hit = (t < v.x) ? 8u * ql : hit;
                    }
                    hit = quad_min(hit);
                    if (hit != 0xffu) {
                        bsel = q0 + (int)hit;
                        break;
                    }
                }
                if (bsel < 0) {
                    topic = sb ? (load_global_u16_at(theta, (start + nnz - 1u) * 2u) & kThetaTopicMask) : (uint32_t)(K - 1);
                } else {
                    // the block's 16 products, four per lane (entries j .. j + 3)
                    const uint32_t j = ((uint32_t)bsel << 4) + 4u * (uint32_t)ql;
                    u32x2 e = u32x2{0u, 0u};
                    float a0, b0, v1, a2, b2, v3;
                    if (sb) {
                        e = load_global2_at(theta, (start + j) * 2u);
                        a0 = COUNT_LO(e.x);
                        b0 = PHI_LO(e.x);
                        v1 = COUNT_HI(e.x) * PHI_HI(e.x);
                        a2 = COUNT_LO(e.y);
                        b2 = PHI_LO(e.y);
                        v3 = COUNT_HI(e.y) * PHI_HI(e.y);
                    } else {
                        const float4 f = *reinterpret_cast<const float4 *>(ph + j);
                        a0 = alpha;
                        b0 = f.x;
                        v1 = alpha * f.y;
                        a2 = alpha;
                        b2 = f.z;
                        v3 = alpha * f.w;
                    }
                    float C0, C1, C2, C3;
                    prefix16x4(a0, b0, v1, a2, b2, v3, C0, C1, C2, C3, ql);
                    const float rb = bsel ? (sb ? rq4 : rq)[bsel - 1] : 0.0f;
                    C0 = rb + C0;
                    C1 = rb + C1;
                    C2 = rb + C2;
                    C3 = rb + C3;
                    const uint32_t lim = sb ? nnz : 0xffffffffu;
                    uint32_t hit = ((j + 3u < lim) & (t < C3)) ? 4u * ql + 3u : 0xffu;
                    hit = ((j + 2u < lim) & (t < C2)) ? 4u * ql + 2u : hit;
                    hit = ((j + 1u < lim) & (t < C1)) ? 4u * ql + 1u : hit;
                    hit = quad_min(((j < lim) & (t < C0)) ? 4u * ql : hit);
                    int msel;
                    if (hit != 0xffu) {
                        msel = (int)hit;
                    } else {
                        const int nvalid = (int)nnz - (bsel << 4);
                        msel = sb ? (nvalid < 16 ? nvalid : 16) - 1 : 15;
                    }
                    // entry msel: lane msel >> 2 of the row, dword (msel >> 1) & 1
                    const int srcl = (qrow << 2) + (msel >> 2);
                    const uint32_t ex = __shfl(e.x, srcl, kWave), ey = __shfl(e.y, srcl, kWave);
                    const uint32_t ew = (msel & 2) ? ey : ex;
                    topic = sb ? (((msel & 1) ? (ew >> 16) : ew) & kThetaTopicMask) : (uint32_t)((bsel << 4) + msel);
                }
            }
            // position r + m's result is in the first lane of row m
            const int src = ((lane - r) & 15) << 2;
            const uint32_t zt = __shfl(topic, src, kWave);
            const float zs = __shfl(SQ, src, kWave);
            if ((unsigned)(lane - r) < 16u) {
                ztop = zt;
                zsq = zs;
            }
        }
        // this lane's token was drawn at position rank of the batch
        const uint32_t mytopic = __shfl(ztop, rank, kWave);
        const float mysq = __shfl(zsq, rank, kWave);
        if (valid) {
            store_global_u16(un.z + base + lane, (uint16_t)mytopic);
            // FP32 term log(fl(SQ / fl(fl(n_d) + fl(K alpha)))), summed in FP64
            ll += (double)logf(mysq / ((float)mnd + (float)md.k_alpha));
            gathered += mnnz;
        }
        wave_sync();  // the next batch's histogram reuses the S-prefix rows
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) ll += __shfl_xor(ll, d, kWave);
    gathered = wave_add_scan_u32(gathered);
    if (lane == 0) a.ll_partial[slot] = ll;
    if (lane == kWave - 1) a.nnz_partial[slot] = (double)gathered;
#undef PHI_LO
#undef PHI_HI
#undef COUNT_LO
#undef COUNT_HI
    if (!kPersistent) break;
    wave_sync();  // LDS is reused by the next slot
    }
}

__global__ void copy_u16_kernel(const uint16_t *src, uint16_t *dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Fixed-order FP64 sums of the per-slot LL partials, per model: stage 1
// sums each packet (a contiguous slot range of one model), stage 2 sums a
// model's packets in order and stores the total at
// history[sweep % kLLHistory][model].  The order depends only on the slot
// lists, so every run adds the same values in the same order.  Inside an
// iteration stage 1 runs as extra workgroups of the count-clear launch and
// stage 2 inside the head-denominator launch (no launches of their own).
// Reference: LDATrainPerplexityReduce1/2 (kernel/lda_train_kernel.cu:273-321),
// which truncated partials to int.
__device__ __forceinline__ void ll_packet_sum(const LLArgs &a, int p) {
    __shared__ double s[256], g[256];
    const Packet pk = a.packets[p];
    double acc = 0.0, nz = 0.0;
    for (int i = pk.first + threadIdx.x; i < pk.last; i += 256) {
        acc += a.partial[i];
        nz += a.nnz_partial[i];
    }
    s[threadIdx.x] = acc;
    g[threadIdx.x] = nz;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s[threadIdx.x] += s[threadIdx.x + w];
            g[threadIdx.x] += g[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.scratch[p] = s[0];
        a.nnz_scratch[p] = g[0];
    }
}

__device__ __forceinline__ void ll_model_sum(const LLArgs &a, int mi) {
    double acc = 0.0, nz = 0.0;
    for (int p = a.model_packets[2 * mi]; p < a.model_packets[2 * mi + 1]; ++p) {
        acc += a.scratch[p];
        nz += a.nnz_scratch[p];
    }
    const size_t h = (size_t)(*a.sweep % kLLHistory) * a.num_models + mi;
    a.history[h] = acc;
    a.nnz_history[h] = nz;
}

__global__ __launch_bounds__(256) void ll_stage1_kernel(LLArgs a) { ll_packet_sum(a, blockIdx.x); }

__global__ __launch_bounds__(64) void ll_stage2_kernel(LLArgs a) {
    if (threadIdx.x == 0 && (int)blockIdx.x < a.num_models) ll_model_sum(a, blockIdx.x);
}

// Joint log-likelihood pieces: sum over nonzero n_wk of
// lgamma(n_wk + beta) - lgamma(beta), one partial per workgroup.
__global__ __launch_bounds__(256) void joint_words_kernel(const int32_t *nwk, int64_t n, double lg_beta, float beta,
                                                          double *partial) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int32_t c = nwk[i];
        if (c) acc += lgamma((double)c + (double)beta) - lg_beta;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// sum over docs and nonzero n_dk of lgamma(n_dk + alpha) - lgamma(alpha);
// n_dk is summed over the topic's consecutive entries, by the lane of its
// last entry (every entry before it holds kThetaMaxCount)
__global__ __launch_bounds__(256) void joint_docs_kernel(const DocMeta *meta, const uint16_t *theta, int32_t num_docs,
                                                         float alpha, double *partial) {
    __shared__ double s[256];
    const double lg_alpha = lgamma((double)alpha);
    double acc = 0.0;
    for (int d = blockIdx.x * 4 + (threadIdx.x >> 6); d < num_docs; d += gridDim.x * 4) {
        const DocMeta m = meta[d];
        for (uint32_t j = threadIdx.x & 63; j < m.nnz; j += 64) {
            const uint32_t e = theta[m.start + j], k = e & kThetaTopicMask;
            if (j + 1 < m.nnz && (theta[m.start + j + 1] & kThetaTopicMask) == k) continue;
            uint32_t n = e >> kThetaCountShift;
            for (uint32_t i = j; i > 0 && (theta[m.start + i - 1] & kThetaTopicMask) == k; --i) n += kThetaMaxCount;
            acc += lgamma((double)n + (double)alpha) - lg_alpha;
        }
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// ------------------------------------------------ allocation-free joint LL
// scratch: the word partials (kJointGrid), one row of kJointGrid doc
// partials per unit (at most kJointUnitsMax units), then the result.  One
// copy per device, held until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns (joint_scratch_ptr).
static __device__ __align__(256) double scratch[kJointGrid * (1 + kJointUnitsMax) + 32];

// One workgroup: the partials in index order, the n_k terms
// lgamma(V beta) - lgamma(n_k + V beta) and the doc-length terms
// lgamma(K alpha) - lgamma(n_d + K alpha) of every unit's docs, each lane's
// share in a fixed order and then a fixed tree; out[0] = the total.
__global__ __launch_bounds__(256) void joint_final_kernel(const double *partial, int32_t num_partials, const int32_t *nk,
                                                          int32_t k, double vb, const UnitDev *units, int32_t unit_begin,
                                                          int32_t unit_end, double ka, double *out) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < num_partials; i += 256) acc += partial[i];
    const double lg_vb = lgamma(vb), lg_ka = lgamma(ka);
    for (int t = threadIdx.x; t < k; t += 256) acc += lg_vb - lgamma((double)nk[t] + vb);
    for (int u = unit_begin; u < unit_end; ++u) {
        const DocMeta *meta = units[u].meta;
        for (int d = threadIdx.x; d < units[u].num_docs; d += 256) acc += lg_ka - lgamma((double)meta[d].nd + ka);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// Theta statistics of one unit's current rows: out[0] += sum_d entries(d),
// out[1] += sum_d n_d * entries(d) (integer sums, one atomic pair per
// workgroup).
__global__ __launch_bounds__(256) void meta_stats_kernel(const DocMeta *meta, int32_t num_docs,
                                                         unsigned long long *out) {
    __shared__ unsigned long long s[2][256];
    unsigned long long n = 0, t = 0;
    for (int d = blockIdx.x * 256 + threadIdx.x; d < num_docs; d += gridDim.x * 256) {
        const DocMeta m = meta[d];
        n += m.nnz;
        t += (unsigned long long)m.nnz * m.nd;
    }
    s[0][threadIdx.x] = n;
    s[1][threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s[0][threadIdx.x] += s[0][threadIdx.x + w];
            s[1][threadIdx.x] += s[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&out[0], s[0][0]);
        atomicAdd(&out[1], s[1][0]);
    }
}

__global__ void copy_kernel(const float4 *src, float4 *dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------------ launchers
static int32_t grid_for(int64_t items, int32_t per_block, int32_t cap) {
    int64_t g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int32_t)g;
}

hipError_t launch_init_z(const InitArgs &a, hipStream_t s) {
    if (a.num_slots <= 0) return hipSuccess;
    hipLaunchKernelGGL(init_z_kernel, dim3(grid_for(a.num_slots, kWavesPerGroup, INT32_MAX)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_add_u32(uint32_t *p, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(add_u32_kernel, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

hipError_t launch_phi_clear(uint32_t *fill, int64_t n_fill, const ModelDev *models, const int2 *rows, int32_t n,
                            hipStream_t s, const LLArgs *ll) {
    const int32_t llp = ll ? std::max(ll->num_packets, 0) : 0;
    if (n_fill <= 0 && n <= 0 && llp == 0) return hipSuccess;
    const int32_t grid = std::max(n_fill > 0 ? grid_for(n_fill, 256, 4096) : 1, n > 0 ? grid_for(n, 1, 4096) : 1);
    LLArgs none{};
    hipLaunchKernelGGL(phi_clear_kernel, dim3(grid + llp), dim3(256), 0, s, fill, n_fill, models, rows, n, grid,
                       llp ? *ll : none);
    return hipGetLastError();
}

hipError_t launch_phi_rebuild(const PhiArgs &a, hipStream_t s) {
    if (a.num_packets <= 0) return hipSuccess;
    hipLaunchKernelGGL(phi_rebuild_kernel, dim3(a.num_packets), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_count_update(const PhiArgs &p, const RareArgs &r, const ThetaArgs &t, hipStream_t s) {
    if (r.mode) return hipErrorInvalidValue;  // the count pass only
    const int64_t grid =
        (int64_t)std::max(t.num_packets, 0) + std::max(p.num_packets, 0) + std::max(r.num_packets, 0);
    if (grid <= 0) return hipSuccess;
    if (grid > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(count_update_kernel, dim3((uint32_t)grid), dim3(256), 0, s, p, r, t);
    return hipGetLastError();
}

hipError_t launch_phi_update(const PhiArgs &p, const RareArgs &r, hipStream_t s) {
    if (r.mode) return hipErrorInvalidValue;  // the count pass only
    const int64_t grid = (int64_t)std::max(p.num_packets, 0) + std::max(r.num_packets, 0);
    if (grid <= 0) return hipSuccess;
    if (grid > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(phi_update_kernel, dim3((uint32_t)grid), dim3(256), 0, s, p, r);
    return hipGetLastError();
}

hipError_t launch_phi_rare(const RareArgs &a, hipStream_t s) {
    if (a.num_packets <= 0) return hipSuccess;
    if (a.mode)
        hipLaunchKernelGGL(phi_rare_kernel<true>, dim3(a.num_packets), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(phi_rare_kernel<false>, dim3(a.num_packets), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_den(const ModelDev *models, int32_t num_models, hipStream_t s, const LLArgs *ll,
                      uint32_t *advance) {
    if (num_models <= 0) return hipSuccess;
    const int32_t grid = grid_for(num_models, 1, 4096);
    if (advance && grid != 1) return hipErrorInvalidValue;  // the advance follows the one workgroup's LL store
    LLArgs none{};
    hipLaunchKernelGGL(den_kernel, dim3(grid), dim3(256), 0, s, models, num_models, ll ? *ll : none, advance);
    return hipGetLastError();
}

hipError_t launch_phi_head(const int32_t *nwk, const float *den, float *head, int32_t num_words, int32_t k,
                           float beta, hipStream_t s) {
    const int64_t n = (int64_t)num_words * k;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(phi_head_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, nwk, den, head, n, k, beta);
    return hipGetLastError();
}

hipError_t launch_theta_rebuild(const ThetaArgs &a, bool alone, int32_t max_groups, hipStream_t s) {
    if (a.num_packets <= 0) return hipSuccess;
    if (alone)
        hipLaunchKernelGGL(theta_rebuild_alone_kernel, dim3(a.num_packets), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(theta_rebuild_kernel,
                           dim3(max_groups > 0 ? std::min(a.num_packets, max_groups) : a.num_packets), dim3(256), 0, s,
                           a);
    return hipGetLastError();
}

hipError_t launch_theta_rebuild_slim(const ThetaArgs &a, int32_t max_groups, hipStream_t s) {
    if (a.num_packets <= 0) return hipSuccess;
    hipLaunchKernelGGL(theta_rebuild_slim_kernel, dim3(std::min(a.num_packets, std::max(max_groups, 1))), dim3(kWave),
                       0, s, a);
    return hipGetLastError();
}

// Row stride of the sampler's S block prefixes for rows of at most
// max_entries theta entries (nb = ceil(max_entries / 16) blocks): a row is
// written at [0, nb] and read, four at a time from multiples of four, below
// 4 * ceil(nb / 4); at least 8 (the rows also hold a batch's 64-word
// histogram).  68 floats while no row has more than 1072 entries.  Every
// 512 bytes of LDS a wave holds costs waves per CU.
int32_t sample_rs_stride(int64_t max_entries) {
    const int64_t nb = (std::max<int64_t>(max_entries, 1) + 15) / 16;
    return (int32_t)std::max<int64_t>(8, 4 * ((nb + 4) / 4));
}

// Row stride of the sixteen 4-lane rows: the 8-lane stride while sixteen
// such rows take at most 1.5 KB, else half of it (the same LDS as the
// eight 8-lane rows).
int32_t sample_rs_stride4(int32_t rs_stride) {
    return rs_stride * 16 * 4 <= 1536 ? rs_stride : std::max(8, rs_stride / 8 * 4);
}

// Dynamic LDS of the sampler: the larger of the two row layouts.
size_t sample_dyn_lds(int32_t rs_stride) {
    return (size_t)std::max(8 * rs_stride, 16 * sample_rs_stride4(rs_stride)) * sizeof(float);
}

// Whether the sampler's LDS (dynamic part: the S-prefix rows) keeps it
// below seven waves per SIMD, in which case the six-wave build, with more
// registers, runs instead.
static bool sample_lds_bound(int32_t rs_stride) {
    static thread_local int32_t last_stride = -1;
    static thread_local bool last = false;
    if (rs_stride != last_stride) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sample_kernel<true, 7>, kWave,
                                                         sample_dyn_lds(rs_stride)) != hipSuccess)
            n = 0;
        last = n > 0 && n < 7 * 4;
        last_stride = rs_stride;
    }
    return last;
}

hipError_t launch_sample(const SampleArgs &a_in, int32_t grid, hipStream_t s) {
    if (a_in.num_slots <= 0) return hipSuccess;
    if (a_in.rs_stride < 8 || a_in.rs_stride % 4) return hipErrorInvalidValue;
    SampleArgs a = a_in;
    a.rs_stride4 = sample_rs_stride4(a.rs_stride);
    const size_t lds = sample_dyn_lds(a.rs_stride);
    const bool six = sample_lds_bound(a.rs_stride);
    if (a.queue) {
        if (six)
            hipLaunchKernelGGL((sample_kernel<true, 6>), dim3(grid), dim3(kWave), lds, s, a);
        else
            hipLaunchKernelGGL((sample_kernel<true, 7>), dim3(grid), dim3(kWave), lds, s, a);
    } else {
        if (six)
            hipLaunchKernelGGL((sample_kernel<false, 6>), dim3(a.num_slots), dim3(kWave), lds, s, a);
        else
            hipLaunchKernelGGL((sample_kernel<false, 7>), dim3(a.num_slots), dim3(kWave), lds, s, a);
    }
    return hipGetLastError();
}

int32_t sample_blocks_per_cu(int32_t rs_stride) {
    int n = 0;
    const size_t lds = sample_dyn_lds(rs_stride);
    const hipError_t e = sample_lds_bound(rs_stride)
                             ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sample_kernel<true, 6>, kWave, lds)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sample_kernel<true, 7>, kWave, lds);
    return e == hipSuccess ? n : 0;
}

// Resident workgroups per CU of the phi rebuild (256 threads, static LDS):
// a launch of at most this many packets per CU runs in one round.
int32_t phi_blocks_per_cu() {
    static int32_t cached = 0;
    static std::once_flag once;
    std::call_once(once, [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, phi_rebuild_kernel, 256, 0) == hipSuccess && n > 0)
            cached = n;
    });
    return cached;
}

hipError_t launch_copy_u16(const uint16_t *src, uint16_t *dst, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_u16_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_ll_reduce(const LLArgs &a, hipStream_t s) {
    if (a.num_packets <= 0) return hipSuccess;
    hipLaunchKernelGGL(ll_stage1_kernel, dim3(a.num_packets), dim3(256), 0, s, a);
    hipLaunchKernelGGL(ll_stage2_kernel, dim3(a.num_models), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_joint_ll_words(const int32_t *nwk, int32_t num_words, int32_t k, double lg_beta, float beta,
                                 double *partial, int32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(joint_words_kernel, dim3(grid), dim3(256), 0, s, nwk, (int64_t)num_words * k, lg_beta, beta,
                       partial);
    return hipGetLastError();
}

hipError_t launch_joint_ll_docs(const DocMeta *meta, const uint16_t *theta, int32_t num_docs, float alpha,
                                double *partial, int32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(joint_docs_kernel, dim3(grid), dim3(256), 0, s, meta, theta, num_docs, alpha, partial);
    return hipGetLastError();
}

// Per-device address of `scratch`, looked up once per device (with that
// device current) by joint_scratch_init.
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

hipError_t joint_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<double *>(p);
    return hipSuccess;
}

double *joint_scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

hipError_t launch_joint_final(const double *partial, int32_t num_partials, const int32_t *nk, int32_t k, double vb,
                              const UnitDev *units, int32_t unit_begin, int32_t unit_end, double ka, double *out,
                              hipStream_t s) {
    hipLaunchKernelGGL(joint_final_kernel, dim3(1), dim3(256), 0, s, partial, num_partials, nk, k, vb, units,
                       unit_begin, unit_end, ka, out);
    return hipGetLastError();
}

hipError_t launch_meta_stats(const DocMeta *meta, int32_t num_docs, unsigned long long *out, hipStream_t s) {
    if (num_docs <= 0) return hipSuccess;
    hipLaunchKernelGGL(meta_stats_kernel, dim3(grid_for(num_docs, 256, 1024)), dim3(256), 0, s, meta, num_docs, out);
    return hipGetLastError();
}

hipError_t launch_copy(const float4 *src, float4 *dst, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n, 256, 65536)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

}  // namespace culda
