// Particle-hole ring (calcChannelT2Diff_phhp, src/calc.cpp:264-313): per mod
// channel, X = V^B + 1/2 V^A T^ph (hp x hp, K = ph) and R = T^ph X
// (ph x hp, K = hp).  Mod blocks are small (at most a few 16-wide tiles per
// side), so the kernel works per wavefront: one wave owns one
// 16-column tile jt of one channel, computes the column tiles X(:, jt) into
// registers and then R(:, jt) = T^ph X(:, jt) from them.  No LDS and no
// barriers: the MFMA accumulator layout of X is reused directly as the B
// operand of the second product by ordering the k-slots to match it.
//
// Arithmetic type CT: double (v_mfma_f64_16x16x4f64) or float
// (v_mfma_f32_16x16x4f32, V^A, V^B and T^ph in FP32);
// R is written in CT (FP32 holds the FP32 sums exactly).  Every load is unconditional: elements
// outside the block read the zero buffer.
#pragma once

#include "device_common.hpp"

struct RingItem {       // one wave's work: mod channel and 16-column tile
    int32_t mod, jt;
};


template <typename CT, int NT>
__device__ __forceinline__ void ring_wave(const DMod &md, int jt, const CT *__restrict__ tph,
                                          const CT *__restrict__ va, const CT *__restrict__ vb,
                                          const double *__restrict__ zero, CT *__restrict__ r) {
    typedef typename Mma<CT>::acc Acc;
    constexpr bool F64 = sizeof(CT) == 8;
    const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    const int PH = md.ph, HP = md.hp;
    const CT *T = tph + md.off_ph;             // PH x HP, [(a,i),(j,b)]
    const CT *A = va + md.off_ph;              // HP x PH, V^A [(i,a),(b,j)]
    const CT *B = vb + md.off_b;               // HP x HP, V^B
    const CT *zc = reinterpret_cast<const CT *>(zero);
    const int col = jt * 16 + m;
    const bool cok = col < HP;

    // k-slot of MFMA step j for lane group kq inside a 16-wide k tile: chosen so
    // that the accumulator register j of a lane holds exactly the X row that the
    // lane supplies as B operand at step j of the second product.
    auto ks = [&](int j) { return F64 ? 4 * j + kq : 4 * kq + j; };

    // X(:, jt) = V^B(:, jt) + 1/2 V^A T(:, jt)
    Acc x[NT];
#pragma unroll
    for (int kt = 0; kt < NT; kt++) {
        const int row = kt * 16 + m;
        const bool rok = row < HP;
        Acc acc = {};
#pragma unroll 1
        for (int pt = 0; pt < NT; pt++) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int k = pt * 16 + ks(j);
                const bool kok = k < PH;
                const CT a = *((rok && kok) ? A + row * PH + k : zc);
                const CT b = *((kok && cok) ? T + k * HP + col : zc);
                acc = mfma16(a, b, acc);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = kt * 16 + Mma<CT>::row(q, lane);
            const CT bv = *((i < HP && cok) ? B + i * HP + col : zc);
            x[kt][q] = (CT)0.5 * acc[q] + bv;
        }
    }
    // R(:, jt) = T X(:, jt); step (kt, j) uses k = kt*16 + ks(j), whose X value
    // is x[kt][j] in this lane.
#pragma unroll 1
    for (int pt = 0; pt < NT; pt++) {
        const int row = pt * 16 + m;
        const bool rok = row < PH;
        Acc acc = {};
#pragma unroll
        for (int kt = 0; kt < NT; kt++) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int k = kt * 16 + ks(j);
                const CT a = *((rok && k < HP) ? T + row * HP + k : zc);
                acc = mfma16(a, x[kt][j], acc);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = pt * 16 + Mma<CT>::row(q, lane);
            if (i < PH && cok) r[md.off_ph + i * HP + col] = acc[q];
        }
    }
}

// LDS-staged ring for mod channels with NT = ceil(max(ph, hp) / 16) <= 5.
// A block of 4 waves takes a group of channels of one NT: four channels of
// 1 tile (a wave each), two of 2 tiles (two waves each), or one of 3, 4 or 5
// tiles (wave w computes the column tiles jt = w, w + 4).  V^A (hp x ph) and
// T^ph (ph x hp) of every channel of the group are staged once in LDS with
// coalesced loads (the per-wave kernel below reads both from global memory
// once per column tile, 16 rows per load instruction); each column tile is
// then computed as in ring_wave, with the same MFMA sequence, so both
// kernels give identical results.  Staged blocks are D x D (D = 16 NT), zero
// outside the channel's block.  Row strides: FP64 odd, so a 16-lane pass of
// b64 reads down a column hits 16 distinct bank pairs; FP32 = 4 mod 32, so
// the 16-byte row reads (4 consecutive k per lane) of 8 lanes and the 4-byte
// column reads of a half-wave hit distinct banks.
template <typename CT, int NT>
struct RingStage {
    static constexpr int D = 16 * NT;
    static constexpr int STR = sizeof(CT) == 8 ? D + 1 : (D + 31) / 32 * 32 + 4;
    static constexpr int G = NT == 1 ? 4 : NT == 2 ? 2 : 1;   // channels per block
    static constexpr size_t CH_BYTES = 2ull * D * STR * sizeof(CT);
    static constexpr size_t BYTES = G * CH_BYTES;
};

struct RingGroup {      // the mod channels of one block of the staged ring (-1: none)
    int32_t mod[4];
};

// Column tile jt of a staged channel (As = V^A, Ts = T^ph in LDS; bv = V^B(:, jt)):
// X(:, jt) = V^B(:, jt) + 1/2 V^A T(:, jt), then R(:, jt) = T X(:, jt), with the
// k-slot of MFMA step j for lane group kq as in ring_wave.
//
// SPLIT (CT = float): both products as BF16x3 split products on
// v_mfma_f32_16x16x16_bf16, three per 16 k instead of four
// v_mfma_f32_16x16x4f32 (about 16 against 34 SIMD cycles each).  Its
// operand lanes hold the values the FP32 steps use (lane group kq: k-slots
// 4 kq .. + 3 of the 16-wide k tile), and its accumulator holds X(4 kq + r,
// col) in register r like the FP32 one, so X is again the B operand of the
// second product as it stands.  Each B operand is split once and used for
// every row tile.
template <typename CT, int NT, int STR, bool SPLIT>
__device__ __forceinline__ void ring_staged_tile(CT (*As)[STR], CT (*Ts)[STR], const DMod &my, int jt,
                                                 const CT (&bv)[NT][4], CT *__restrict__ r) {
    typedef typename Mma<CT>::acc Acc;
    constexpr bool F64 = sizeof(CT) == 8;
    const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    const int PH = my.ph, HP = my.hp, col = jt * 16 + m;
    const bool cok = col < HP;
    if constexpr (SPLIT) {
        static_assert(!F64, "split products from FP32 operands");
        Split4 tb[NT];   // B of the first product: Ts(pt*16 + 4 kq + e, col)
#pragma unroll
        for (int pt = 0; pt < NT; pt++)
            tb[pt] = split4(Ts[pt * 16 + 4 * kq][col], Ts[pt * 16 + 4 * kq + 1][col], Ts[pt * 16 + 4 * kq + 2][col],
                            Ts[pt * 16 + 4 * kq + 3][col]);
        Split4 xs[NT];   // X(:, jt), split: B of the second product
#pragma unroll
        for (int kt = 0; kt < NT; kt++) {
            flt4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int pt = 0; pt < NT; pt++) {
                const flt4 a4 = *reinterpret_cast<const flt4 *>(&As[kt * 16 + m][pt * 16 + 4 * kq]);
                acc = mfma_bf16x3_k16(split4(a4[0], a4[1], a4[2], a4[3]), tb[pt], acc);
            }
            float xv[4];
#pragma unroll
            for (int q = 0; q < 4; q++) xv[q] = 0.5f * acc[q] + bv[kt][q];
            xs[kt] = split4(xv[0], xv[1], xv[2], xv[3]);
        }
#pragma unroll
        for (int pt = 0; pt < NT; pt++) {
            flt4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < NT; kt++) {
                const flt4 a4 = *reinterpret_cast<const flt4 *>(&Ts[pt * 16 + m][kt * 16 + 4 * kq]);
                acc = mfma_bf16x3_k16(split4(a4[0], a4[1], a4[2], a4[3]), xs[kt], acc);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = pt * 16 + Mma<CT>::row(q, lane);
                if (i < PH && cok) r[my.off_ph + i * HP + col] = acc[q];
            }
        }
        return;
    }
    Acc x[NT];
#pragma unroll
    for (int kt = 0; kt < NT; kt++) {
        Acc acc = {};
#pragma unroll
        for (int pt = 0; pt < NT; pt++) {
            if constexpr (F64) {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc = mfma16(As[kt * 16 + m][pt * 16 + 4 * j + kq], Ts[pt * 16 + 4 * j + kq][col], acc);
            } else {
                const flt4 a4 = *reinterpret_cast<const flt4 *>(&As[kt * 16 + m][pt * 16 + 4 * kq]);
#pragma unroll
                for (int j = 0; j < 4; j++) acc = mfma16(a4[j], Ts[pt * 16 + 4 * kq + j][col], acc);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) x[kt][q] = (CT)0.5 * acc[q] + bv[kt][q];
    }
#pragma unroll
    for (int pt = 0; pt < NT; pt++) {
        Acc acc = {};
#pragma unroll
        for (int kt = 0; kt < NT; kt++) {
            if constexpr (F64) {
#pragma unroll
                for (int j = 0; j < 4; j++) acc = mfma16(Ts[pt * 16 + m][kt * 16 + 4 * j + kq], x[kt][j], acc);
            } else {
                const flt4 a4 = *reinterpret_cast<const flt4 *>(&Ts[pt * 16 + m][kt * 16 + 4 * kq]);
#pragma unroll
                for (int j = 0; j < 4; j++) acc = mfma16(a4[j], x[kt][j], acc);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = pt * 16 + Mma<CT>::row(q, lane);
            if (i < PH && cok) r[my.off_ph + i * HP + col] = acc[q];
        }
    }
}

template <typename CT, int NT, bool SPLIT, int NW>
__device__ __forceinline__ void ring_staged_block(const RingGroup &grp, const DMod *__restrict__ mods,
                                                  const CT *__restrict__ tph, const CT *__restrict__ va,
                                                  const CT *__restrict__ vb, const double *__restrict__ zero,
                                                  CT *__restrict__ r, unsigned char *lds) {
    typedef RingStage<CT, NT> S;
    constexpr int D = S::D, STR = S::STR, G = S::G;
    // the wave index is uniform across the wave: kept in scalar registers, so
    // are the channel descriptor and the column-tile loop below
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, m = lane & 15;
    const CT *zc = reinterpret_cast<const CT *>(zero);
    DMod md[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int k = grp.mod[g];
        md[g] = k >= 0 ? mods[k] : DMod{0, 0, 0, 0, 0, 0};
    }
    // this wave's channel and first column tile
    const int g0 = NT == 1 ? wave : NT == 2 ? wave >> 1 : 0;
    const int jt0 = NT == 1 ? 0 : NT == 2 ? (wave & 1) : wave;
    DMod my = md[0];
#pragma unroll
    for (int g = 1; g < G; g++)
        if (g0 == g) my = md[g];
    const int HP = my.hp;

    // V^B(:, jt0), the initial value of X(:, jt0): issued first, consumed after the first product
    CT bv[NT][4];
    {
        const CT *B = vb + my.off_b;
        const int col = jt0 * 16 + m;
#pragma unroll
        for (int kt = 0; kt < NT; kt++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = kt * 16 + Mma<CT>::row(q, lane);
                bv[kt][q] = *((i < HP && col < HP) ? B + i * HP + col : zc);
            }
    }

    // staging in rounds of up to 8 elements per thread (every load of a round
    // is issued before its LDS stores): element e = thread + TB q of the group
    // (TB threads) lies in channel e / D^2 = q / (D^2 / TB), a constant of the
    // unrolled loops
    constexpr int TB = 64 * NW, CR = D * D / TB;   // threads; rounds per channel
    constexpr int PER = G * CR;
    static_assert(D * D % TB == 0, "whole staging rounds");
    constexpr int RND = PER < 8 ? PER : 8;
#pragma unroll
    for (int q0 = 0; q0 < PER; q0 += RND) {
        CT sa[RND], st[RND];
#pragma unroll
        for (int u = 0; u < RND; u++) {
            const int q = q0 + u;
            if (q < PER) {
                const int g = q / CR, rem = threadIdx.x + (q - g * CR) * TB;
                const int row = rem / D, c = rem - row * D;
                const DMod &h = md[g];
                sa[u] = *((row < h.hp && c < h.ph) ? va + h.off_ph + row * h.ph + c : zc);
                st[u] = *((row < h.ph && c < h.hp) ? tph + h.off_ph + row * h.hp + c : zc);
            }
        }
#pragma unroll
        for (int u = 0; u < RND; u++) {
            const int q = q0 + u;
            if (q < PER) {
                const int g = q / CR, rem = threadIdx.x + (q - g * CR) * TB;
                const int row = rem / D, c = rem - row * D;
                CT(*As)[STR] = reinterpret_cast<CT(*)[STR]>(lds + g * S::CH_BYTES);
                CT(*Ts)[STR] = reinterpret_cast<CT(*)[STR]>(lds + g * S::CH_BYTES + sizeof(CT) * D * STR);
                As[row][c] = sa[u];
                Ts[row][c] = st[u];
            }
        }
    }
    __syncthreads();

    CT(*As)[STR] = reinterpret_cast<CT(*)[STR]>(lds + g0 * S::CH_BYTES);
    CT(*Ts)[STR] = reinterpret_cast<CT(*)[STR]>(lds + g0 * S::CH_BYTES + sizeof(CT) * D * STR);
    if (jt0 * 16 < HP) ring_staged_tile<CT, NT, STR, SPLIT>(As, Ts, my, jt0, bv, r);
    if (NT == 5 && NW == 4 && wave == 0 && 4 * 16 < HP) {   // the fifth column tile (four waves)
        const CT *B = vb + my.off_b;
        const int col = 4 * 16 + m;
#pragma unroll
        for (int kt = 0; kt < NT; kt++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = kt * 16 + Mma<CT>::row(q, lane);
                bv[kt][q] = *((i < HP && col < HP) ? B + i * HP + col : zc);
            }
        ring_staged_tile<CT, NT, STR, SPLIT>(As, Ts, my, 4, bv, r);
    }
}

// Groups of channels with LO..HI tiles (largest first), one group per block
// of NW waves.  SPLIT: BF16x3 split products (FP32 iterations).  Channels of
// 5 tiles run with NW = 5, a wave per column tile (with four waves, wave 0
// would take the fifth tile as well, and the block would wait for it).
template <typename CT, int LO, int HI, bool SPLIT = false, int NW = 4>
__global__ __launch_bounds__(64 * NW) void k_ring_staged(const RingGroup *__restrict__ groups,
                                                          const DMod *__restrict__ mods,
                                                          const CT *__restrict__ tph,
                                                          const CT *__restrict__ va, const CT *__restrict__ vb,
                                                          const double *__restrict__ zero,
                                                          CT *__restrict__ r,
                                                          const LoopCtl *__restrict__ ctl,
                                                          const LoopCtl *__restrict__ pctl) {
    if (ctl->done) return;
    extern __shared__ __align__(16) unsigned char ring_lds[];
    const RingGroup grp = groups[blockIdx.x];
    if (pctl) {   // a batch: skip a group whose channels all belong to stopped solves
        bool all = true;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (grp.mod[k] >= 0 && !block_solve_stopped(pctl, mods[grp.mod[k]].prob)) all = false;
        if (all) return;
    }
    const DMod m0 = mods[grp.mod[0]];
    const int nt = (max(m0.ph, m0.hp) + 15) / 16;
    static_for<LO, HI + 1>([&](auto N) {
        constexpr int NT = decltype(N)::value;
        if (nt == NT) ring_staged_block<CT, NT, SPLIT, NW>(grp, mods, tph, va, vb, zero, r, ring_lds);
    });
}

// Items are sorted by tile count, largest first; wave w of block b takes item
// 4b + w.  A channel with NT = ceil(max(ph, hp) / 16) has ceil(hp / 16) items.
// WIDE instances take the items with NT = 6..8, the others NT = 1..5 (the
// wide register footprint would lower the occupancy of every item).
template <typename CT, bool WIDE>
__global__ __launch_bounds__(kThreads) void k_ring(const RingItem *__restrict__ items, int nitems,
                                                   const DMod *__restrict__ mods,
                                                   const CT *__restrict__ tph,
                                                   const CT *__restrict__ va, const CT *__restrict__ vb,
                                                   const double *__restrict__ zero,
                                                   CT *__restrict__ r,
                                                   const LoopCtl *__restrict__ ctl,
                                                   const LoopCtl *__restrict__ pctl) {
    if (ctl->done) return;
    const int it = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (it >= nitems) return;
    const RingItem w = items[it];
    const DMod md = mods[w.mod];
    if (wave_solve_stopped(pctl, md.prob)) return;   // a stopped solve of a batch
    const int nt = (max(md.ph, md.hp) + 15) / 16;
    if (WIDE) {
        switch (nt) {
            case 6: ring_wave<CT, 6>(md, w.jt, tph, va, vb, zero, r); break;
            case 7: ring_wave<CT, 7>(md, w.jt, tph, va, vb, zero, r); break;
            default: ring_wave<CT, 8>(md, w.jt, tph, va, vb, zero, r); break;
        }
    } else {
        switch (nt) {
            case 1: ring_wave<CT, 1>(md, w.jt, tph, va, vb, zero, r); break;
            case 2: ring_wave<CT, 2>(md, w.jt, tph, va, vb, zero, r); break;
            case 3: ring_wave<CT, 3>(md, w.jt, tph, va, vb, zero, r); break;
            case 4: ring_wave<CT, 4>(md, w.jt, tph, va, vb, zero, r); break;
            default: ring_wave<CT, 5>(md, w.jt, tph, va, vb, zero, r); break;
        }
    }
}
