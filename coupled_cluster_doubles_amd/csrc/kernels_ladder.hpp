// Particle-particle / hole-hole ladders and the hhhh quadratic term
// (calcChannelT2Diff_pphh, src/calc.cpp:132-203).
//
// With t and V antisymmetric in (a,b) and (c,d), the reference's ordered-pair
// ladder 1/2 sum_{cd} V[ab][cd] t[cd][ij] equals sum_{c<d} V[ab][cd] t[cd][ij],
// and D[(b,a)] = -D[(a,b)], D[(a,a)] = 0.  Both ladder kernels therefore run on
// the packed pairs a<b: a quarter of the V^pppp bytes and flops.  The
// contractions use v_mfma_f64_16x16x4f64 (layout in device_common.hpp).
#pragma once

#include "device_common.hpp"

constexpr int KC = 32;                 // k per chunk of k_ladder_x (k-slot q of step s: k = 8q + s)
constexpr int NTMAX = 5;               // 16-column tiles: hh <= 80
constexpr int LROWS = 64;              // packed rows per ladder block (16 per wave)

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
// columns, split over k segments of kseg_pairs(NT) packed pairs (tile.aux =
// segment): segment s writes its partial sum to slab s of the channel's D_p,
// and the update adds the slabs in segment order.  The k range of a block is
// one sequence of KCH-wide chunks: for segment 0 first the hh columns of T X,
// then its packed pairs of V t.  In a chunk, k-slot q of MFMA step s is
// k = 8 (s/2) + 2q + s%2: the four lanes of a row read 64 contiguous bytes per
// 16-byte load, so a load instruction touches half cache lines whose other
// halves the next instruction reads (rows are 128-byte aligned).  Software
// pipeline: while the MFMAs of chunk c run, the A values of chunk c+1 (HBM)
// and its B tile (L2-resident amplitudes / X) are loaded into registers; the
// B tile then goes to the other LDS buffer, one barrier per chunk.
__device__ __forceinline__ int kslot(int q, int s) { return 8 * (s >> 1) + 2 * q + (s & 1); }

template <int NT>
struct LadderShape {
    static constexpr int KCH = NT <= 2 ? 64 : 32;   // k per pipeline stage
    static constexpr int KSEG = NT <= 2 ? 256 : 64; // packed pairs per k segment
};

template <int NT>
__global__ __launch_bounds__(kThreads) void k_ladder_v(const DTile *__restrict__ tiles,
                                                       const DChan *__restrict__ chans,
                                                       const double *__restrict__ vp,
                                                       const double *__restrict__ tas,
                                                       const double *__restrict__ x,
                                                       double *__restrict__ dp) {
    constexpr int KCH = LadderShape<NT>::KCH, KSEG = LadderShape<NT>::KSEG, G = KCH / 4;
    constexpr int W = NT * 16, STR = W + 1, BPT = (KCH * W + kThreads - 1) / kThreads;
    __shared__ double Bs[2][KCH][STR];
    const DTile tl = tiles[blockIdx.x];
    const DChan ch = chans[tl.chan];
    const int H = ch.hh, npk = ch.npk, ldv = ch.ldv, seg = tl.aux;
    const int kbeg = seg * KSEG, kend = min(npk, kbeg + KSEG);
    const int nx = seg == 0 ? (H + KCH - 1) / KCH : 0, nchunk = nx + (kend - kbeg + KCH - 1) / KCH;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    const int lrow = wave * 16 + m;
    const bool rok = lrow < tl.nrows;
    const int64_t u = tl.row0 + lrow;
    const double *at = tas + ch.off_tp + u * H;
    const double *av = vp + ch.off_vp + u * ldv;

    auto load_a = [&](int c, double (&a)[G]) {
        if (c < nx) {
#pragma unroll
            for (int s = 0; s < G; s++) {
                const int k = c * KCH + kslot(kq, s);
                a[s] = (rok && k < H) ? 0.5 * at[k] : 0.0;
            }
        } else {
            const int kc0 = kbeg + (c - nx) * KCH;
#pragma unroll
            for (int h = 0; h < G / 2; h++) {
                const int k = kc0 + 8 * h + 2 * kq;  // even; ldv is even, so k+1 < ldv too
                if (rok && k < kend && k < ldv) {
                    const dbl2 v = *reinterpret_cast<const dbl2 *>(av + k);
                    a[2 * h] = v.x;
                    a[2 * h + 1] = v.y;
                } else {
                    a[2 * h] = 0.0;
                    a[2 * h + 1] = 0.0;
                }
            }
        }
    };
    auto load_b = [&](int c, double (&b)[BPT]) {
        const double *B = c < nx ? x + ch.off_hhhh : tas + ch.off_tp;
        const int K = c < nx ? H : kend, k0 = c < nx ? c * KCH : kbeg + (c - nx) * KCH;
#pragma unroll
        for (int q = 0; q < BPT; q++) {
            const int idx = threadIdx.x + kThreads * q, kk = idx / W, col = idx - kk * W, k = k0 + kk;
            b[q] = (idx < KCH * W && k < K && col < H) ? B[(int64_t)k * H + col] : 0.0;
        }
    };
    auto store_b = [&](int buf, const double (&b)[BPT]) {
#pragma unroll
        for (int q = 0; q < BPT; q++) {
            const int idx = threadIdx.x + kThreads * q, kk = idx / W, col = idx - kk * W;
            if (idx < KCH * W) Bs[buf][kk][col] = b[q];
        }
    };

    dbl4 acc[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) acc[n] = dbl4{0.0, 0.0, 0.0, 0.0};
    double a[G], an[G] = {}, b[BPT];
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
        for (int s = 0; s < G; s++) {
#pragma unroll
            for (int n = 0; n < NT; n++) acc[n] = mfma16(a[s], Bs[cb][kslot(kq, s)][n * 16 + m], acc[n]);
        }
        if (more) store_b(cb ^ 1, b);
        __syncthreads();
#pragma unroll
        for (int s = 0; s < G; s++) a[s] = an[s];
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

// Number of k segments of a channel's pppp ladder.
inline int ladder_segments(int npk, int nt) {
    const int kseg = nt <= 2 ? LadderShape<1>::KSEG : LadderShape<3>::KSEG;
    return npk > 0 ? (npk + kseg - 1) / kseg : 1;
}
