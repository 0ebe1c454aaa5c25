// Particle-hole ring (calcChannelT2Diff_phhp, src/calc.cpp:264-313):
// per mod channel, X = V^B + 1/2 V^A T^ph (hp x hp, K = ph) and
// R = T^ph X (ph x hp, K = hp), both as v_mfma_f64_16x16x4f64 tiles.
#pragma once

#include "device_common.hpp"

// One mod channel per block.  X lives in LDS (row stride xs_stride doubles,
// at least hp rounded up to 16); T^ph, V^A and V^B are read from global
// memory (a channel's blocks are at most a few tens of KB and stay in L1/L2).
// Within a 32-wide k chunk, k-slot q of MFMA step s is k = 8q + s, so each
// lane reads 8 consecutive elements of its A row.
__global__ __launch_bounds__(kThreads) void k_ring(const DMod *__restrict__ mods, int xs_stride,
                                                   const double *__restrict__ tph,
                                                   const double *__restrict__ va,
                                                   const double *__restrict__ vb,
                                                   double *__restrict__ r) {
    extern __shared__ double xs[];
    const DMod m = mods[blockIdx.x];
    const int PH = m.ph, HP = m.hp;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, mm = lane & 15, kq = lane >> 4;
    const double *T = tph + m.off_ph;  // PH x HP
    const double *A = va + m.off_ph;   // HP x PH
    const double *B = vb + m.off_b;    // HP x HP
    const int nth = (HP + 15) / 16, ntp = (PH + 15) / 16;

    // X tiles
    for (int t = wave; t < nth * nth; t += kThreads / 64) {
        const int kt = t / nth, jt = t - kt * nth;
        const int row = kt * 16 + mm, col = jt * 16 + mm;
        dbl4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < PH; k0 += 32) {
            double a[8], b[8];
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const int p = k0 + kq * 8 + s;
                const bool ok = p < PH;
                a[s] = (ok && row < HP) ? A[row * PH + p] : 0.0;
                b[s] = (ok && col < HP) ? T[p * HP + col] : 0.0;
            }
#pragma unroll
            for (int s = 0; s < 8; s++) acc = mfma16(a[s], b[s], acc);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = kt * 16 + 4 * q + kq;
            if (i < HP && col < HP) xs[i * xs_stride + col] = 0.5 * acc[q] + B[i * HP + col];
        }
    }
    __syncthreads();
    // R tiles
    for (int t = wave; t < ntp * nth; t += kThreads / 64) {
        const int pt = t / nth, jt = t - pt * nth;
        const int row = pt * 16 + mm, col = jt * 16 + mm;
        dbl4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < HP; k0 += 32) {
            double a[8], b[8];
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const int k = k0 + kq * 8 + s;
                const bool ok = k < HP;
                a[s] = (ok && row < PH) ? T[row * HP + k] : 0.0;
                b[s] = (ok && col < HP) ? xs[k * xs_stride + col] : 0.0;
            }
#pragma unroll
            for (int s = 0; s < 8; s++) acc = mfma16(a[s], b[s], acc);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = pt * 16 + 4 * q + kq;
            if (i < PH && col < HP) r[m.off_ph + i * HP + col] = acc[q];
        }
    }
}
