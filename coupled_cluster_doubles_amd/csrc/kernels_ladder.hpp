// Particle-particle / hole-hole ladders and the hhhh quadratic term
// (calcChannelT2Diff_pphh, src/calc.cpp:132-203).
#pragma once

#include "device_common.hpp"

//
// With t and V antisymmetric in (a,b) and (c,d), the reference's ordered-pair
// ladder 1/2 sum_{cd} V[ab][cd] t[cd][ij] equals sum_{c<d} V[ab][cd] t[cd][ij],
// and D[(b,a)] = -D[(a,b)], D[(a,a)] = 0.  Both ladder kernels therefore run on
// the packed pairs a<b: a quarter of the V^pppp bytes and flops.  The
// contractions use v_mfma_f64_16x16x4f64: A is 16x4 (lane l: row l%16, k-slot
// l/16), B is 4x16 (lane l: k-slot l/16, column l%16), C is 16x16 with
// register q of lane l holding row 4q + l/16, column l%16.  Within a 32-wide k chunk, k-slot q
// of step s is k = 8q + s, so each lane reads 8 consecutive A elements.

constexpr int KC = 32;                 // k rows per LDS stage
constexpr int NTMAX = 5;               // 16-column tiles: hh <= 80
constexpr int BSTR = NTMAX * 16 + 1;   // LDS row stride (doubles)
constexpr int LROWS = 64;              // packed rows per ladder block (16 per wave)

// Bs[kk][c] = B[(k0+kk)*ldb + c] for c < nt*16, zero where k >= K or c >= ncols.
__device__ __forceinline__ void stage_b(double (*Bs)[BSTR], const double *__restrict__ B, int ldb,
                                        int k0, int K, int ncols, int nt) {
    const int w = nt * 16;
    for (int idx = threadIdx.x; idx < KC * w; idx += kThreads) {
        const int kk = idx / w, c = idx - kk * w, k = k0 + kk;
        Bs[kk][c] = (k < K && c < ncols) ? B[(int64_t)k * ldb + c] : 0.0;
    }
}

__device__ __forceinline__ void mfma_chunk(const double (&a)[8], double (*Bs)[BSTR], int kq, int m, int nt,
                                           dbl4 (&acc)[NTMAX]) {
#pragma unroll
    for (int s = 0; s < 8; s++) {
#pragma unroll
        for (int n = 0; n < NTMAX; n++) {
            if (n < nt) acc[n] = mfma16(a[s], Bs[kq * 8 + s][n * 16 + m], acc[n]);
        }
    }
}

// X = V^hhhh + sum_{c<d} V^hhpp[kl][cd] tas[cd][ij]; one wave per 16x16 tile of X.
__global__ __launch_bounds__(kThreads) void k_ladder_x(const DTile *__restrict__ tiles, int ntiles,
                                                       const DChan *__restrict__ chans,
                                                       const double *__restrict__ vtp,
                                                       const double *__restrict__ tas,
                                                       const double *__restrict__ vh,
                                                       double *__restrict__ x) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    const int tix = blockIdx.x * (kThreads / 64) + wave;
    if (tix >= ntiles) return;
    const DTile tl = tiles[tix];
    const DChan ch = chans[tl.chan];
    const int H = ch.hh, npk = ch.npk, kl0 = tl.row0, ij0 = tl.aux;
    const bool aok = kl0 + m < H, bok = ij0 + m < H;
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < npk; k0 += KC) {
        double a[8], b[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int w = k0 + kq * 8 + s;
            const bool ok = w < npk;
            a[s] = (ok && aok) ? vtp[ch.off_tp + (int64_t)w * H + kl0 + m] : 0.0;
            b[s] = (ok && bok) ? tas[ch.off_tp + (int64_t)w * H + ij0 + m] : 0.0;
        }
#pragma unroll
        for (int s = 0; s < 8; s++) acc = mfma16(a[s], b[s], acc);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int row = kl0 + q * 4 + kq, col = ij0 + m;
        if (row < H && col < H) {
            const int64_t o = ch.off_hhhh + (int64_t)row * H + col;
            x[o] = vh[o] + acc[q];
        }
    }
}

// D_p[u][ij] = 1/2 sum_kl tas[u][kl] X[kl][ij] + sum_w Vp[u][w] tas[w][ij]
// for 64 packed rows u of one channel per block (16 per wave), all hh columns.
__global__ __launch_bounds__(kThreads) void k_ladder_v(const DTile *__restrict__ tiles,
                                                       const DChan *__restrict__ chans,
                                                       const double *__restrict__ vp,
                                                       const double *__restrict__ tas,
                                                       const double *__restrict__ x,
                                                       double *__restrict__ dp) {
    __shared__ double Bs[KC][BSTR];
    const DTile tl = tiles[blockIdx.x];
    const DChan ch = chans[tl.chan];
    const int H = ch.hh, npk = ch.npk, ldv = ch.ldv;
    const int nt = (H + 15) / 16;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    const int lrow = wave * 16 + m;
    const bool rok = lrow < tl.nrows;
    const int64_t u = tl.row0 + lrow;
    dbl4 acc[NTMAX];
#pragma unroll
    for (int n = 0; n < NTMAX; n++) acc[n] = dbl4{0.0, 0.0, 0.0, 0.0};

    // 1/2 T X  (K = hh, ordered kl)
    const double *at = tas + ch.off_tp + u * H;
    for (int k0 = 0; k0 < H; k0 += KC) {
        __syncthreads();
        stage_b(Bs, x + ch.off_hhhh, H, k0, H, H, nt);
        __syncthreads();
        double a[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int k = k0 + kq * 8 + s;
            a[s] = (rok && k < H) ? 0.5 * at[k] : 0.0;
        }
        mfma_chunk(a, Bs, kq, m, nt, acc);
    }
    // V^pppp t  (K = npk packed pairs; rows padded with zeros to ldv)
    const double *av = vp + ch.off_vp + u * ldv;
    for (int k0 = 0; k0 < npk; k0 += KC) {
        __syncthreads();
        stage_b(Bs, tas + ch.off_tp, H, k0, npk, H, nt);
        __syncthreads();
        double a[8];
        const int kb = k0 + kq * 8;
        if (rok && kb < ldv) {       // ldv and kb are multiples of 8: whole 64-byte group in range
            const dbl2 *p2 = reinterpret_cast<const dbl2 *>(av + kb);
#pragma unroll
            for (int h = 0; h < 4; h++) {
                const dbl2 v = p2[h];
                a[2 * h] = v.x;
                a[2 * h + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int s = 0; s < 8; s++) a[s] = 0.0;
        }
        mfma_chunk(a, Bs, kq, m, nt, acc);
    }
#pragma unroll
    for (int n = 0; n < NTMAX; n++) {
        if (n >= nt) continue;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int row = wave * 16 + q * 4 + kq, col = n * 16 + m;
            if (row < tl.nrows && col < H) dp[ch.off_tp + (int64_t)(tl.row0 + row) * H + col] = acc[n][q];
        }
    }
}

