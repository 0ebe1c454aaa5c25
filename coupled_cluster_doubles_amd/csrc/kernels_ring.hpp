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
// (v_mfma_f32_16x16x4f32, V^A and V^B stored in FP32, T^ph rounded on load);
// R is written in FP64 either way.  Every load is unconditional: elements
// outside the block read the zero buffer.
#pragma once

#include "device_common.hpp"

struct RingItem {       // one wave's work: mod channel and 16-column tile
    int32_t mod, jt;
};


template <typename CT, int NT>
__device__ __forceinline__ void ring_wave(const DMod &md, int jt, const double *__restrict__ tph,
                                          const CT *__restrict__ va, const CT *__restrict__ vb,
                                          const double *__restrict__ zero, double *__restrict__ r) {
    typedef typename Mma<CT>::acc Acc;
    constexpr bool F64 = sizeof(CT) == 8;
    const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    const int PH = md.ph, HP = md.hp;
    const double *T = tph + md.off_ph;         // PH x HP, [(a,i),(j,b)]
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
                const CT b = (CT) * ((kok && cok) ? T + k * HP + col : zero);
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
                const CT a = (CT) * ((rok && k < HP) ? T + row * HP + k : zero);
                acc = mfma16(a, x[kt][j], acc);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = pt * 16 + Mma<CT>::row(q, lane);
            if (i < PH && cok) r[md.off_ph + i * HP + col] = (double)acc[q];
        }
    }
}

// Items are sorted by tile count, largest first; wave w of block b takes item
// 4b + w.  A channel with NT = ceil(max(ph, hp) / 16) has ceil(hp / 16) items.
// WIDE instances take the items with NT = 6..8, the others NT = 1..5 (the
// wide register footprint would lower the occupancy of every item).
template <typename CT, bool WIDE>
__global__ __launch_bounds__(kThreads) void k_ring(const RingItem *__restrict__ items, int nitems,
                                                   const DMod *__restrict__ mods,
                                                   const double *__restrict__ tph,
                                                   const CT *__restrict__ va, const CT *__restrict__ vb,
                                                   const double *__restrict__ zero,
                                                   double *__restrict__ r,
                                                   const LoopCtl *__restrict__ ctl) {
    if (ctl->done) return;
    const int it = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (it >= nitems) return;
    const RingItem w = items[it];
    const DMod md = mods[w.mod];
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
