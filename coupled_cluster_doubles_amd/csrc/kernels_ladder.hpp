// Particle-particle / hole-hole ladders and the hhhh quadratic term
// (calcChannelT2Diff_pphh, src/calc.cpp:132-203).
//
// With t and V antisymmetric in (a,b) and (c,d), the reference's ordered-pair
// ladder 1/2 sum_{cd} V[ab][cd] t[cd][ij] equals sum_{c<d} V[ab][cd] t[cd][ij],
// and D[(b,a)] = -D[(a,b)], D[(a,a)] = 0.  Both ladder kernels therefore run on
// the packed pairs a<b: a quarter of the V^pppp bytes and flops.  The
// contractions use v_mfma_f64_16x16x4f64 (layout in device_common.hpp).
// Within a 32-wide k chunk, k-slot q of MFMA step s is k = 8q + s, so each
// lane reads 8 consecutive elements of its A row.
#pragma once

#include "device_common.hpp"

constexpr int KC = 32;                 // k per pipeline stage
constexpr int NTMAX = 5;               // 16-column tiles: hh <= 80
constexpr int LROWS = 64;              // packed rows per ladder block (16 per wave)
constexpr int KSEG = 8 * KC;           // packed pairs per k segment of the pppp ladder

// X = V^hhhh + sum_{c<d} V^hhpp[kl][cd] tas[cd][ij], one 16x16 tile of X per
// block; the four waves take every fourth k chunk and their partial tiles are
// added in wave order through LDS.
__global__ __launch_bounds__(kThreads) void k_ladder_x(const DTile *__restrict__ tiles,
                                                       const DChan *__restrict__ chans,
                                                       const double *__restrict__ vtp,
                                                       const double *__restrict__ tas,
                                                       const double *__restrict__ vh,
                                                       double *__restrict__ x) {
    __shared__ double part[kThreads / 64][64][4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    const DTile tl = tiles[blockIdx.x];
    const DChan ch = chans[tl.chan];
    const int H = ch.hh, npk = ch.npk, kl0 = tl.row0, ij0 = tl.aux;
    const bool aok = kl0 + m < H, bok = ij0 + m < H;
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = wave * KC; k0 < npk; k0 += KC * (kThreads / 64)) {
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
    for (int q = 0; q < 4; q++) part[wave][lane][q] = acc[q];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const double s = ((part[0][lane][q] + part[1][lane][q]) + part[2][lane][q]) + part[3][lane][q];
            const int row = kl0 + q * 4 + kq, col = ij0 + m;
            if (row < H && col < H) {
                const int64_t o = ch.off_hhhh + (int64_t)row * H + col;
                x[o] = vh[o] + s;
            }
        }
    }
}

// D_p[u][ij] = 1/2 sum_kl tas[u][kl] X[kl][ij] + sum_w Vp[u][w] tas[w][ij]
// for 64 packed rows u of one channel per block (16 per wave) and NT*16 >= hh
// columns, split over k segments of KSEG packed pairs (tile.aux = segment):
// segment s writes its partial sum to slab s of the channel's D_p, and the
// update adds the slabs in segment order.  The k range of a block is one
// sequence of 32-wide chunks: for segment 0 first the hh columns of T X, then
// its packed pairs of V t.  Software pipeline: while the
// MFMAs of chunk c run, the A values of chunk c+1 (HBM) and its B tile
// (L2-resident amplitudes / X) are loaded into registers; the B tile then
// goes to the other LDS buffer, one barrier per chunk.
template <int NT>
__global__ __launch_bounds__(kThreads) void k_ladder_v(const DTile *__restrict__ tiles,
                                                       const DChan *__restrict__ chans,
                                                       const double *__restrict__ vp,
                                                       const double *__restrict__ tas,
                                                       const double *__restrict__ x,
                                                       double *__restrict__ dp) {
    constexpr int W = NT * 16, STR = W + 1, BPT = (KC * W + kThreads - 1) / kThreads;
    __shared__ double Bs[2][KC][STR];
    const DTile tl = tiles[blockIdx.x];
    const DChan ch = chans[tl.chan];
    const int H = ch.hh, npk = ch.npk, ldv = ch.ldv, seg = tl.aux;
    const int kbeg = seg * KSEG, kend = min(npk, kbeg + KSEG);
    const int nx = seg == 0 ? (H + KC - 1) / KC : 0, nchunk = nx + (kend - kbeg + KC - 1) / KC;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    const int lrow = wave * 16 + m;
    const bool rok = lrow < tl.nrows;
    const int64_t u = tl.row0 + lrow;
    const double *at = tas + ch.off_tp + u * H;
    const double *av = vp + ch.off_vp + u * ldv;

    auto load_a = [&](int c, double (&a)[8]) {
        if (c < nx) {
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const int k = c * KC + kq * 8 + s;
                a[s] = (rok && k < H) ? 0.5 * at[k] : 0.0;
            }
        } else {
            const int kb = kbeg + (c - nx) * KC + kq * 8;
            if (rok && kb < ldv && kb < kend) {  // ldv, kb, kend - kbeg multiples of 8: whole 64-byte group
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
        }
    };
    auto load_b = [&](int c, double (&b)[BPT]) {
        const double *B = c < nx ? x + ch.off_hhhh : tas + ch.off_tp;
        const int K = c < nx ? H : kend, k0 = c < nx ? c * KC : kbeg + (c - nx) * KC;
#pragma unroll
        for (int q = 0; q < BPT; q++) {
            const int idx = threadIdx.x + kThreads * q, kk = idx / W, col = idx - kk * W, k = k0 + kk;
            b[q] = (idx < KC * W && k < K && col < H) ? B[(int64_t)k * H + col] : 0.0;
        }
    };
    auto store_b = [&](int buf, const double (&b)[BPT]) {
#pragma unroll
        for (int q = 0; q < BPT; q++) {
            const int idx = threadIdx.x + kThreads * q, kk = idx / W, col = idx - kk * W;
            if (idx < KC * W) Bs[buf][kk][col] = b[q];
        }
    };

    dbl4 acc[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) acc[n] = dbl4{0.0, 0.0, 0.0, 0.0};
    double a[8], an[8] = {}, b[BPT];
    load_a(0, a);
    load_b(0, b);
    store_b(0, b);
    __syncthreads();
    for (int c = 0; c < nchunk; c++) {
        const bool more = c + 1 < nchunk;
        if (more) {
            load_a(c + 1, an);
            load_b(c + 1, b);
        }
        const int cb = c & 1;
#pragma unroll
        for (int s = 0; s < 8; s++) {
#pragma unroll
            for (int n = 0; n < NT; n++) acc[n] = mfma16(a[s], Bs[cb][kq * 8 + s][n * 16 + m], acc[n]);
        }
        if (more) store_b(cb ^ 1, b);
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 8; s++) a[s] = an[s];
    }
#pragma unroll
    for (int n = 0; n < NT; n++) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int row = wave * 16 + q * 4 + kq, col = n * 16 + m;
            if (row < tl.nrows && col < H)
                dp[ch.off_dp + (int64_t)seg * npk * H + (int64_t)(tl.row0 + row) * H + col] = acc[n][q];
        }
    }
}
