// MI355X (gfx950) implementation of the CCD amplitude iteration.
//
// Replaces, for one GPU, the hot path of justinlietz/coupled_cluster_doubles:
//   calcT2Diff_pphh   src/calc.cpp:132-231   -> k_ladder_x + k_ladder_d
//   calcT2Diff_phhp   src/calc.cpp:264-339   -> k_ring
//   calcT2Diff_h2pp_h1, calcT2Diff_p2_p1hh   -> k_onebody (chi, psi) + k_update
//   8 x mapFromDistributed, 5 x mapToDistributed (src/Chain.hpp:710-802)
//                                            -> index arithmetic inside k_update
//   calcT2_pphh       src/calc.cpp:368-459   -> k_update + k_sum_partials
//   MBPT2Corr         src/MBPT2Corr.cpp:6-38 -> k_mbpt2 + k_sum_partials
//   loop              src/main.cpp:491-682   -> ccd_solve
//
// HBM layout: every amplitude-shaped array (T ping-pong, D, W) is the
// reference pphh Chain layout; T^ph and R the phhp layout; V^hhpp is kept
// transposed (pphh layout, "vt") so the update reads it with the same index
// as T.  V^pppp, V^hhhh, V^B are stored for the active channels only.
// The one-body intermediates are diagonal (chi_i, psi_a), so the reference's
// 3-body chains and their four remaps reduce to segmented sums of
// W = V o T and two in-channel swaps; the four ring gathers and the T^ph
// scatter are computed from two small position tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "ccd_amd.h"
#include "internal.hpp"

// ------------------------------------------------------------------ errors

namespace ccd {
static thread_local std::string g_err;
int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
void clear_error() { g_err.clear(); }
}  // namespace ccd

extern "C" const char *ccd_last_error(void) { return ccd::g_err.c_str(); }

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ccd::fail(CCD_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                           \
    } while (0)

// ------------------------------------------------------------------ types

struct DChan {          // active direct channel (pp > 0 and hh > 0)
    int64_t off_t;      // block offset in the pphh layout
    int64_t off_vp;     // packed V^pppp block: npk x ldv, rows/cols = pairs a<b
    int64_t off_tp;     // packed-row blocks (npk x hh): antisymmetrised T, D_p, V^hhpp rows
    int64_t off_hhhh;   // hh x hh block of V^hhhh / X
    int32_t pp, hh;
    int32_t pp_off, hh_off;  // first pair in the pp / hh pair lists
    int32_t npk, ldv;        // number of pairs a<b; leading dimension (multiple of 8)
    int32_t pr_off, pad;     // first entry in the packed-row lists
};

struct DMod {           // active mod channel (ph > 0 and hp > 0)
    int64_t off_ph;     // block offset in the phhp layout (also V^A, hp x ph)
    int64_t off_b;      // offset of its hp x hp block of V^B
    int32_t ph, hp;
    int32_t ph_off, hp_off;
};

struct DTile {          // a run of rows of one active direct channel
    int32_t chan, row0, nrows, pad;
};

struct DSeg {           // strided run of pphh elements for a one-body sum
    int64_t e0;
    int32_t stride, count;
};

struct Gen {            // on-device interaction
    int kind;
    int nmax;
    double g;
    const int32_t *qn;
    const double *table;
};

constexpr int kThreads = 256;

// ------------------------------------------------------------------ device helpers

__device__ __forceinline__ double block_sum(double v, double *red) {
    // fixed-order reduction: wave butterfly, then the four wave sums in order
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
    return s;  // valid on thread 0
}

// Antisymmetrised <pq||rs> generated from quantum numbers; bitwise equal to
// the host model (pairing: src/PairingSPBasis.cpp:127-141; electron gas:
// src/ElectronGasSPBasis.cpp:419-471 with the direct term tabulated).
__device__ double gen_tbme(const Gen &G, int p, int q, int r, int s) {
    const int32_t *a = G.qn + 5 * p, *b = G.qn + 5 * q, *c = G.qn + 5 * r, *d = G.qn + 5 * s;
    if (G.kind == CCD_GEN_PAIRING) {
        double v = 0.0;
        if (a[0] == b[0] && c[0] == d[0] && a[1] == -b[1] && -c[1] == d[1]) {
            v = -0.5 * G.g;
            if (a[1] == -c[1]) v = -v;
        }
        return v;
    }
    const bool cons = a[0] + b[0] == c[0] + d[0] && a[1] + b[1] == c[1] + d[1] && a[2] + b[2] == c[2] + d[2];
    if (!cons) return 0.0;
    const int m2 = 2 * G.nmax, w = 4 * G.nmax + 1;
    double d1 = 0.0, d2 = 0.0;
    if (a[3] == c[3] && b[3] == d[3])
        d1 = G.table[((a[0] - c[0] + m2) * w + (a[1] - c[1] + m2)) * w + (a[2] - c[2] + m2)];
    if (a[3] == d[3] && b[3] == c[3])
        d2 = G.table[((a[0] - d[0] + m2) * w + (a[1] - d[1] + m2)) * w + (a[2] - d[2] + m2)];
    return d1 - d2;
}

// ------------------------------------------------------------------ generation kernels

// Packed V^pppp rows of one tile: Vp[u][w] = <a_u b_u||c_w d_w> over pairs a<b,
// zero in the padding columns w >= npk.
__global__ __launch_bounds__(kThreads) void k_gen_vp(const DTile *__restrict__ tiles,
                                                     const DChan *__restrict__ chans,
                                                     const int2 *__restrict__ ppl,
                                                     const int32_t *__restrict__ prow, Gen G,
                                                     double *__restrict__ vp) {
    const DTile tl = tiles[blockIdx.x];
    const DChan ch = chans[tl.chan];
    const int ldv = ch.ldv, npk = ch.npk;
    const int64_t n = (int64_t)tl.nrows * ldv;
    for (int64_t l = threadIdx.x; l < n; l += kThreads) {
        const int u = tl.row0 + (int)(l / ldv), w = (int)(l % ldv);
        double v = 0.0;
        if (w < npk) {
            const int2 ab = ppl[ch.pp_off + prow[ch.pr_off + u]];
            const int2 cd = ppl[ch.pp_off + prow[ch.pr_off + w]];
            v = gen_tbme(G, ab.x, ab.y, cd.x, cd.y);
        }
        vp[ch.off_vp + (int64_t)u * ldv + w] = v;
    }
}

// Packed rows of V^hhpp^T: vtp[u][kl] = vt[(a_u,b_u)][kl] (tile = packed rows).
__global__ __launch_bounds__(kThreads) void k_pack_vt(const DTile *__restrict__ tiles,
                                                      const DChan *__restrict__ chans,
                                                      const int32_t *__restrict__ prow,
                                                      const double *__restrict__ vt,
                                                      double *__restrict__ vtp) {
    const DTile tl = tiles[blockIdx.x];
    const DChan ch = chans[tl.chan];
    const int H = ch.hh;
    for (int l = threadIdx.x; l < tl.nrows * H; l += kThreads) {
        const int u = tl.row0 + l / H, kl = l % H;
        vtp[ch.off_tp + (int64_t)u * H + kl] = vt[ch.off_t + (int64_t)prow[ch.pr_off + u] * H + kl];
    }
}

// Antisymmetrised amplitudes on the packed rows: tas[u][ij] = (t[(a,b)][ij] - t[(b,a)][ij]) / 2.
__global__ __launch_bounds__(kThreads) void k_antisym(const DTile *__restrict__ tiles,
                                                      const DChan *__restrict__ chans,
                                                      const int32_t *__restrict__ prow,
                                                      const int32_t *__restrict__ prow_sw,
                                                      const double *__restrict__ t,
                                                      double *__restrict__ tas) {
    const DTile tl = tiles[blockIdx.x];
    const DChan ch = chans[tl.chan];
    const int H = ch.hh;
    for (int l = threadIdx.x; l < tl.nrows * H; l += kThreads) {
        const int u = tl.row0 + l / H, ij = l % H;
        const double x = t[ch.off_t + (int64_t)prow[ch.pr_off + u] * H + ij];
        const double y = t[ch.off_t + (int64_t)prow_sw[ch.pr_off + u] * H + ij];
        tas[ch.off_tp + (int64_t)u * H + ij] = 0.5 * (x - y);
    }
}

// V^hhpp transposed into the pphh layout, [(a,b),(i,j)] = <ij||ab>, and V^hhhh.
__global__ __launch_bounds__(kThreads) void k_gen_direct_small(const DChan *__restrict__ chans,
                                                               const int2 *__restrict__ ppl,
                                                               const int2 *__restrict__ hhl, Gen G,
                                                               double *__restrict__ vt,
                                                               double *__restrict__ vh) {
    const DChan ch = chans[blockIdx.x];
    const int P = ch.pp, H = ch.hh;
    for (int l = threadIdx.x; l < P * H; l += kThreads) {
        const int AB = l / H, IJ = l - AB * H;
        const int2 ab = ppl[ch.pp_off + AB], ij = hhl[ch.hh_off + IJ];
        vt[ch.off_t + l] = gen_tbme(G, ij.x, ij.y, ab.x, ab.y);
    }
    for (int l = threadIdx.x; l < H * H; l += kThreads) {
        const int K = l / H, J = l - K * H;
        const int2 kl = hhl[ch.hh_off + K], ij = hhl[ch.hh_off + J];
        vh[ch.off_hhhh + l] = gen_tbme(G, kl.x, kl.y, ij.x, ij.y);
    }
}

// V^A [(i,a),(b,j)] = <ij||ab> (hp x ph) and V^B [(i,b),(j,a)] = <ia||bj> (hp x hp)
// of one mod channel (src/load.cpp:275-318).
__global__ __launch_bounds__(kThreads) void k_gen_mod(const DMod *__restrict__ mods,
                                                      const int2 *__restrict__ phl,
                                                      const int2 *__restrict__ hpl, Gen G,
                                                      double *__restrict__ va,
                                                      double *__restrict__ vb) {
    const DMod m = mods[blockIdx.x];
    const int PH = m.ph, HP = m.hp;
    for (int l = threadIdx.x; l < HP * PH; l += kThreads) {
        const int k = l / PH, c = l - k * PH;
        const int2 ia = hpl[m.hp_off + k], bj = phl[m.ph_off + c];
        va[m.off_ph + l] = gen_tbme(G, ia.x, bj.y, ia.y, bj.x);
    }
    for (int l = threadIdx.x; l < HP * HP; l += kThreads) {
        const int k = l / HP, c = l - k * HP;
        const int2 ib = hpl[m.hp_off + k], ja = hpl[m.hp_off + c];
        vb[m.off_b + l] = gen_tbme(G, ib.x, ja.y, ib.y, ja.x);
    }
}

// Caller V^hhpp (hh x pp) -> pphh layout.
__global__ __launch_bounds__(kThreads) void k_transpose_hhpp(const DChan *__restrict__ chans,
                                                             const double *__restrict__ vhhpp,
                                                             double *__restrict__ vt) {
    const DChan ch = chans[blockIdx.x];
    const int P = ch.pp, H = ch.hh;
    for (int l = threadIdx.x; l < P * H; l += kThreads) {
        const int AB = l / H, IJ = l - AB * H;
        vt[ch.off_t + l] = vhhpp[ch.off_t + (int64_t)IJ * P + AB];
    }
}

// ------------------------------------------------------------------ ring (calc.cpp:264-313)

// One mod channel per block: X = V^B + 1/2 V^A T^ph (hp x hp, in LDS),
// R = T^ph X (ph x hp).
__global__ __launch_bounds__(kThreads) void k_ring(const DMod *__restrict__ mods,
                                                   const double *__restrict__ tph,
                                                   const double *__restrict__ va,
                                                   const double *__restrict__ vb,
                                                   double *__restrict__ r) {
    extern __shared__ double sx[];
    const DMod m = mods[blockIdx.x];
    const int PH = m.ph, HP = m.hp;
    const double *T = tph + m.off_ph;
    const double *A = va + m.off_ph;
    const double *B = vb + m.off_b;
    for (int o = threadIdx.x; o < HP * HP; o += kThreads) {
        const int k = o / HP, j = o - k * HP;
        double acc = 0.0;
        for (int p = 0; p < PH; p++) acc += A[k * PH + p] * T[p * HP + j];
        sx[o] = 0.5 * acc + B[o];
    }
    __syncthreads();
    for (int o = threadIdx.x; o < PH * HP; o += kThreads) {
        const int p = o / HP, j = o - p * HP;
        double acc = 0.0;
        for (int k = 0; k < HP; k++) acc += T[p * HP + k] * sx[k * HP + j];
        r[m.off_ph + o] = acc;
    }
}

// ------------------------------------------------------------------ ladders (calc.cpp:132-203)
//
// With t and V antisymmetric in (a,b) and (c,d), the reference's ordered-pair
// ladder 1/2 sum_{cd} V[ab][cd] t[cd][ij] equals sum_{c<d} V[ab][cd] t[cd][ij],
// and D[(b,a)] = -D[(a,b)], D[(a,a)] = 0.  Both ladder kernels therefore run on
// the packed pairs a<b: a quarter of the V^pppp bytes and flops.  The
// contractions use v_mfma_f64_16x16x4f64: A is 16x4 (lane l: row l%16, k-slot
// l/16), B is 4x16 (lane l: k-slot l/16, column l%16), C is 16x16 with
// register q of lane l holding row 4q + l/16, column l%16.  Within a 32-wide k chunk, k-slot q
// of step s is k = 8q + s, so each lane reads 8 consecutive A elements.

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));

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
            if (n < nt) acc[n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], Bs[kq * 8 + s][n * 16 + m], acc[n], 0, 0, 0);
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
    const int H = ch.hh, npk = ch.npk, kl0 = tl.row0, ij0 = tl.pad;
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
        for (int s = 0; s < 8; s++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
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

// ------------------------------------------------------------------ update (calc.cpp:368-421)

struct UpdArgs {
    const DTile *tiles;
    const DChan *chans;
    const int2 *ppl, *hhl;
    const int32_t *pp_swap, *hh_swap;  // index of (b,a) / (j,i) within the channel
    const int32_t *pkc;                // per pp pair: -1 if a==b, else (packed row << 1) | (a > b)
    const int32_t *rowbase;            // [(a-nh)*nh + i] -> phhp offset of row (a,i)
    const int32_t *colidx;             // [j*np + (b-nh)]  -> column of (j,b)
    const double *eps, *chi, *psi;
    const double *vt, *d, *r, *t_old;
    double *t_new, *tph, *w;
    double *partials;                  // per block sum of W
    double mixing;
    int nh, np;
};

__global__ __launch_bounds__(kThreads) void k_update(UpdArgs a) {
    __shared__ double red[kThreads / 64];
    const DTile tl = a.tiles[blockIdx.x];
    const DChan ch = a.chans[tl.chan];
    const int H = ch.hh;
    const int nh = a.nh, np = a.np;
    double wsum = 0.0;
    for (int l = threadIdx.x; l < tl.nrows * H; l += kThreads) {
        const int AB = tl.row0 + l / H, IJ = l % H;
        const int64_t e = ch.off_t + (int64_t)AB * H + IJ;
        const int2 ab = a.ppl[ch.pp_off + AB], ij = a.hhl[ch.hh_off + IJ];
        const int ia = ab.x, ib = ab.y, ii = ij.x, jj = ij.y;
        const int pa = ia - nh, pb = ib - nh;
        const int g1 = a.rowbase[pa * nh + ii] + a.colidx[jj * np + pb];  // (a,i),(j,b)
        const int g2 = a.rowbase[pa * nh + jj] + a.colidx[ii * np + pb];  // (a,j),(i,b)
        const int g3 = a.rowbase[pb * nh + ii] + a.colidx[jj * np + pa];  // (b,i),(j,a)
        const int g4 = a.rowbase[pb * nh + jj] + a.colidx[ii * np + pa];  // (b,j),(i,a)
        const double t = a.t_old[e];
        const double t_ji = a.t_old[ch.off_t + (int64_t)AB * H + a.hh_swap[ch.hh_off + IJ]];
        const double t_ba = a.t_old[ch.off_t + (int64_t)a.pp_swap[ch.pp_off + AB] * H + IJ];
        const double v = a.vt[e];
        double value = v;
        const int code = a.pkc[ch.pp_off + AB];
        double dv = 0.0;
        if (code >= 0) {
            dv = a.d[ch.off_tp + (int64_t)(code >> 1) * H + IJ];
            if (code & 1) dv = -dv;
        }
        value += dv;
        value += a.r[g1];
        value -= a.r[g2];
        value -= a.r[g3];
        value += a.r[g4];
        value += (0.5 * a.chi[ii]) * t_ji;   // h2pp_h1 gathered as jab_i
        value -= (0.5 * a.chi[jj]) * t;      // h2pp_h1 gathered as iab_j
        value += (-0.5 * a.psi[pb]) * t;     // p2_p1hh gathered as b_aij
        value -= (-0.5 * a.psi[pa]) * t_ba;  // p2_p1hh gathered as a_bij
        value = value / (a.eps[ii] + a.eps[jj] - a.eps[ia] - a.eps[ib]);
        value = a.mixing * value + (1. - a.mixing) * t;
        a.t_new[e] = value;
        a.tph[g1] = value;                   // mapToDistributed(ai_jb)
        const double wv = v * value;
        a.w[e] = wv;
        wsum += wv;
    }
    const double s = block_sum(wsum, red);
    if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
}

// MBPT2 terms v^2 / denominator, per-block partial sums.
__global__ __launch_bounds__(kThreads) void k_mbpt2(const DTile *__restrict__ tiles,
                                                    const DChan *__restrict__ chans,
                                                    const int2 *__restrict__ ppl,
                                                    const int2 *__restrict__ hhl,
                                                    const double *__restrict__ eps,
                                                    const double *__restrict__ vt,
                                                    double *__restrict__ partials) {
    __shared__ double red[kThreads / 64];
    const DTile tl = tiles[blockIdx.x];
    const DChan ch = chans[tl.chan];
    const int H = ch.hh;
    double s = 0.0;
    for (int l = threadIdx.x; l < tl.nrows * H; l += kThreads) {
        const int AB = tl.row0 + l / H, IJ = l % H;
        const int2 ab = ppl[ch.pp_off + AB], ij = hhl[ch.hh_off + IJ];
        const double v = vt[ch.off_t + (int64_t)AB * H + IJ];
        const double den = eps[ij.x] + eps[ij.y] - eps[ab.x] - eps[ab.y];
        s += v * v / den;
    }
    const double b = block_sum(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

// out = scale * sum(partials[0..n)) in a fixed order (one block).
__global__ __launch_bounds__(kThreads) void k_sum_partials(const double *__restrict__ partials,
                                                           int n, double scale,
                                                           double *__restrict__ out) {
    __shared__ double red[kThreads / 64];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += kThreads) s += partials[k];
    const double b = block_sum(s, red);
    if (threadIdx.x == 0) out[0] = scale * b;
}

// chi_i and psi_a: block k sums the segments seg_ptr[k]..seg_ptr[k+1] of W.
__global__ __launch_bounds__(kThreads) void k_onebody(const int32_t *__restrict__ seg_ptr,
                                                      const DSeg *__restrict__ segs,
                                                      const double *__restrict__ w,
                                                      double *__restrict__ out) {
    __shared__ double red[kThreads / 64];
    const int s0 = seg_ptr[blockIdx.x], s1 = seg_ptr[blockIdx.x + 1];
    double s = 0.0;
    for (int k = s0 + threadIdx.x; k < s1; k += kThreads) {
        const DSeg sg = segs[k];
        double t = 0.0;
        for (int q = 0; q < sg.count; q++) t += w[sg.e0 + (int64_t)q * sg.stride];
        s += t;
    }
    const double b = block_sum(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = b;
}

// Scatter T (pphh) into T^ph (phhp): the standalone remap entry point.
__global__ __launch_bounds__(kThreads) void k_remap_ph(UpdArgs a) {
    const DTile tl = a.tiles[blockIdx.x];
    const DChan ch = a.chans[tl.chan];
    const int H = ch.hh, nh = a.nh, np = a.np;
    for (int l = threadIdx.x; l < tl.nrows * H; l += kThreads) {
        const int AB = tl.row0 + l / H, IJ = l % H;
        const int2 ab = a.ppl[ch.pp_off + AB], ij = a.hhl[ch.hh_off + IJ];
        const int g1 = a.rowbase[(ab.x - nh) * nh + ij.x] + a.colidx[ij.y * np + (ab.y - nh)];
        a.tph[g1] = a.t_old[ch.off_t + (int64_t)AB * H + IJ];
    }
}

// D in the ordered pphh layout from D_p (for the ladder entry point).
__global__ __launch_bounds__(kThreads) void k_expand_d(UpdArgs a) {
    const DTile tl = a.tiles[blockIdx.x];
    const DChan ch = a.chans[tl.chan];
    const int H = ch.hh;
    for (int l = threadIdx.x; l < tl.nrows * H; l += kThreads) {
        const int AB = tl.row0 + l / H, IJ = l % H;
        const int code = a.pkc[ch.pp_off + AB];
        double dv = 0.0;
        if (code >= 0) {
            dv = a.d[ch.off_tp + (int64_t)(code >> 1) * H + IJ];
            if (code & 1) dv = -dv;
        }
        a.t_new[ch.off_t + (int64_t)AB * H + IJ] = dv;
    }
}

// W = V o T (for the one-body entry point).
__global__ __launch_bounds__(kThreads) void k_hadamard(const double *__restrict__ vt,
                                                       const double *__restrict__ t,
                                                       double *__restrict__ w, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * kThreads)
        w[e] = vt[e] * t[e];
}

// ------------------------------------------------------------------ context

struct ccd_ctx {
    int device = 0;
    int precision = CCD_PREC_FP64;
    hipStream_t stream = nullptr;
    int nsp = 0, nh = 0, np = 0;
    int64_t nt2 = 0, nph = 0;
    int64_t bytes = 0;
    std::vector<DChan> chans;
    std::vector<DMod> mods;
    std::vector<DTile> ltiles, utiles, xtiles;
    std::vector<double> hbuf;
    // device
    DChan *d_chans = nullptr;
    DMod *d_mods = nullptr;
    DTile *d_ltiles = nullptr, *d_utiles = nullptr, *d_xtiles = nullptr;
    int2 *d_ppl = nullptr, *d_hhl = nullptr, *d_phl = nullptr, *d_hpl = nullptr;
    int32_t *d_pp_swap = nullptr, *d_hh_swap = nullptr, *d_rowbase = nullptr, *d_colidx = nullptr;
    int32_t *d_prow = nullptr, *d_prow_sw = nullptr, *d_pkc = nullptr;
    int32_t *d_seg_ptr = nullptr;
    DSeg *d_segs = nullptr;
    double *d_eps = nullptr, *d_onebody = nullptr;  // chi [nh] then psi [np]
    double *d_t[2] = {nullptr, nullptr};
    int cur = 0;
    double *d_d = nullptr, *d_tph = nullptr, *d_r = nullptr, *d_w = nullptr, *d_x = nullptr;
    double *d_vt = nullptr, *d_vh = nullptr, *d_vp = nullptr, *d_va = nullptr, *d_vb = nullptr;
    double *d_vtp = nullptr, *d_tas = nullptr;
    double *d_partials = nullptr, *d_scalar = nullptr;
    double *h_scalar = nullptr;  // pinned
    int32_t *d_qn = nullptr;
    double *d_table = nullptr;
    int64_t n_vp = 0, n_tp = 0, n_hhhh = 0, n_b = 0, n_vp_alg = 0;
    size_t ring_lds_max = 0;
    double flops = 0, flops_pppp = 0;
    std::vector<void *> allocs;
};

namespace {

template <typename T>
int dalloc(ccd_ctx *c, T **p, size_t count) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess)
        return ccd::fail(CCD_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    c->allocs.push_back((void *)*p);
    c->bytes += (int64_t)bytes;
    return CCD_OK;
}

template <typename T>
int upload(ccd_ctx *c, T **p, const std::vector<T> &h) {
    int rc = dalloc(c, p, h.size());
    if (rc) return rc;
    if (!h.empty()) HIP_TRY(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return CCD_OK;
}

void free_ctx(ccd_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->h_scalar) (void)hipHostFree(c->h_scalar);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ccd::fail(CCD_E_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return CCD_OK;
}

UpdArgs upd_args(ccd_ctx *c, const double *t_old, double *t_new, double mixing) {
    UpdArgs a{};
    a.tiles = c->d_utiles; a.chans = c->d_chans; a.ppl = c->d_ppl; a.hhl = c->d_hhl;
    a.pp_swap = c->d_pp_swap; a.hh_swap = c->d_hh_swap; a.pkc = c->d_pkc; a.rowbase = c->d_rowbase;
    a.colidx = c->d_colidx; a.eps = c->d_eps; a.chi = c->d_onebody; a.psi = c->d_onebody + c->nh;
    a.vt = c->d_vt; a.d = c->d_d; a.r = c->d_r; a.t_old = t_old; a.t_new = t_new;
    a.tph = c->d_tph; a.w = c->d_w; a.partials = c->d_partials; a.mixing = mixing;
    a.nh = c->nh; a.np = c->np;
    return a;
}

// Phases of one iteration; ev (optional) brackets each phase.
int launch_ring(ccd_ctx *c) {
    if (c->mods.empty()) return CCD_OK;
    hipLaunchKernelGGL(k_ring, dim3((unsigned)c->mods.size()), dim3(kThreads), c->ring_lds_max,
                       c->stream, c->d_mods, c->d_tph, c->d_va, c->d_vb, c->d_r);
    return check_launch("k_ring");
}

int launch_ladder_x(ccd_ctx *c, const double *t) {
    if (c->ltiles.empty()) return CCD_OK;
    hipLaunchKernelGGL(k_antisym, dim3((unsigned)c->ltiles.size()), dim3(kThreads), 0, c->stream, c->d_ltiles,
                       c->d_chans, c->d_prow, c->d_prow_sw, t, c->d_tas);
    int rc = check_launch("k_antisym");
    if (rc || c->xtiles.empty()) return rc;
    const unsigned nb = (unsigned)((c->xtiles.size() + 3) / 4);
    hipLaunchKernelGGL(k_ladder_x, dim3(nb), dim3(kThreads), 0, c->stream, c->d_xtiles, (int)c->xtiles.size(),
                       c->d_chans, c->d_vtp, c->d_tas, c->d_vh, c->d_x);
    return check_launch("k_ladder_x");
}

int launch_ladder_d(ccd_ctx *c, const double *t) {
    (void)t;
    if (c->ltiles.empty()) return CCD_OK;
    hipLaunchKernelGGL(k_ladder_v, dim3((unsigned)c->ltiles.size()), dim3(kThreads), 0, c->stream,
                       c->d_ltiles, c->d_chans, c->d_vp, c->d_tas, c->d_x, c->d_d);
    return check_launch("k_ladder_v");
}

int launch_update(ccd_ctx *c, const double *t_old, double *t_new, double mixing) {
    if (c->utiles.empty()) return CCD_OK;
    UpdArgs a = upd_args(c, t_old, t_new, mixing);
    hipLaunchKernelGGL(k_update, dim3((unsigned)c->utiles.size()), dim3(kThreads), 0, c->stream, a);
    return check_launch("k_update");
}

int launch_reductions(ccd_ctx *c) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, c->stream, c->d_partials,
                       (int)c->utiles.size(), 0.25, c->d_scalar);
    int rc = check_launch("k_sum_partials");
    if (rc) return rc;
    hipLaunchKernelGGL(k_onebody, dim3((unsigned)(c->nh + c->np)), dim3(kThreads), 0, c->stream,
                       c->d_seg_ptr, c->d_segs, c->d_w, c->d_onebody);
    return check_launch("k_onebody");
}

int run_iteration(ccd_ctx *c, double mixing, hipEvent_t *ev) {
    const double *t_old = c->d_t[c->cur];
    double *t_new = c->d_t[1 - c->cur];
    int rc;
    if (ev) HIP_TRY(hipEventRecord(ev[0], c->stream));
    if ((rc = launch_ring(c))) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[1], c->stream));
    if ((rc = launch_ladder_x(c, t_old))) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[2], c->stream));
    if ((rc = launch_ladder_d(c, t_old))) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[3], c->stream));
    if ((rc = launch_update(c, t_old, t_new, mixing))) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[4], c->stream));
    if ((rc = launch_reductions(c))) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[5], c->stream));
    c->cur = 1 - c->cur;
    return CCD_OK;
}

int read_scalar(ccd_ctx *c, double *out) {
    HIP_TRY(hipMemcpyAsync(c->h_scalar, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *out = *c->h_scalar;
    return CCD_OK;
}

int validate(const ccd_problem *p) {
    if (!p) return ccd::fail(CCD_E_INVALID, "null problem");
    if (p->n_sp <= 0 || p->n_holes < 0 || p->n_holes > p->n_sp || p->n_chan < 0 || p->n_mod < 0)
        return ccd::fail(CCD_E_INVALID, "bad sizes (n_sp=%d n_holes=%d n_chan=%d n_mod=%d)", p->n_sp,
                         p->n_holes, p->n_chan, p->n_mod);
    if (!p->eps || (p->n_chan && (!p->pp_dim || !p->hh_dim)) || (p->n_mod && (!p->ph_dim || !p->hp_dim)))
        return ccd::fail(CCD_E_INVALID, "missing arrays in problem");
    if (p->gen_kind == CCD_GEN_NONE) {
        if (!p->v_hhhh || !p->v_hhpp || !p->v_pppp || !p->v_modA || !p->v_modB)
            return ccd::fail(CCD_E_INVALID, "gen_kind NONE needs all five V chains");
    } else if (p->gen_kind == CCD_GEN_PAIRING || p->gen_kind == CCD_GEN_ELECTRON_GAS) {
        if (!p->qn) return ccd::fail(CCD_E_INVALID, "generator needs qn");
        if (p->gen_kind == CCD_GEN_ELECTRON_GAS && (!p->gen_table || p->gen_nmax < 0))
            return ccd::fail(CCD_E_INVALID, "electron-gas generator needs gen_table and gen_nmax");
    } else {
        return ccd::fail(CCD_E_INVALID, "unknown gen_kind %d", p->gen_kind);
    }
    return CCD_OK;
}

int build(ccd_ctx *c, const ccd_problem *p) {
    const int nsp = p->n_sp, nh = p->n_holes, np = nsp - nh;
    c->nsp = nsp; c->nh = nh; c->np = np;
    // --- pair lists and channel tables (host) ---
    int64_t npp = 0, nhh = 0, nphp = 0, nhpp = 0;
    for (int k = 0; k < p->n_chan; k++) {
        if (p->pp_dim[k] < 0 || p->hh_dim[k] < 0) return ccd::fail(CCD_E_INVALID, "negative dim");
        npp += p->pp_dim[k]; nhh += p->hh_dim[k];
    }
    for (int k = 0; k < p->n_mod; k++) {
        if (p->ph_dim[k] < 0 || p->hp_dim[k] < 0) return ccd::fail(CCD_E_INVALID, "negative dim");
        nphp += p->ph_dim[k]; nhpp += p->hp_dim[k];
    }
    if ((npp && !p->pp_pairs) || (nhh && !p->hh_pairs) || (nphp && !p->ph_pairs) || (nhpp && !p->hp_pairs))
        return ccd::fail(CCD_E_INVALID, "missing pair lists");
    auto pair_ok = [&](const int32_t *l, int64_t n, bool p_hole, bool q_hole) {
        for (int64_t k = 0; k < n; k++) {
            int a = l[2 * k], b = l[2 * k + 1];
            if (a < 0 || a >= nsp || b < 0 || b >= nsp) return false;
            if ((a < nh) != p_hole || (b < nh) != q_hole) return false;
        }
        return true;
    };
    if (!pair_ok(p->pp_pairs, npp, false, false) || !pair_ok(p->hh_pairs, nhh, true, true) ||
        !pair_ok(p->ph_pairs, nphp, false, true) || !pair_ok(p->hp_pairs, nhpp, true, false))
        return ccd::fail(CCD_E_INVALID, "pair list entry out of range or in the wrong class");

    std::vector<int2> ppl(npp), hhl(nhh), phl(nphp), hpl(nhpp);
    for (int64_t k = 0; k < npp; k++) ppl[k] = make_int2(p->pp_pairs[2 * k], p->pp_pairs[2 * k + 1]);
    for (int64_t k = 0; k < nhh; k++) hhl[k] = make_int2(p->hh_pairs[2 * k], p->hh_pairs[2 * k + 1]);
    for (int64_t k = 0; k < nphp; k++) phl[k] = make_int2(p->ph_pairs[2 * k], p->ph_pairs[2 * k + 1]);
    for (int64_t k = 0; k < nhpp; k++) hpl[k] = make_int2(p->hp_pairs[2 * k], p->hp_pairs[2 * k + 1]);

    // direct channels
    std::vector<int32_t> pp_swap(npp, -1), hh_swap(nhh, -1), pkc(npp, -1);
    std::vector<int32_t> prow, prow_sw;
    int64_t off_t = 0, off_vp = 0, off_tp = 0, off_hhhh = 0, ipp = 0, ihh = 0;
    c->flops = 0; c->flops_pppp = 0; c->n_vp_alg = 0;
    std::unordered_map<int64_t, int32_t> pos;
    for (int k = 0; k < p->n_chan; k++) {
        const int P = p->pp_dim[k], H = p->hh_dim[k];
        if (P > 0 && H > 0) {
            // swaps within the channel
            pos.clear();
            for (int q = 0; q < P; q++) pos[(int64_t)ppl[ipp + q].x * nsp + ppl[ipp + q].y] = q;
            for (int q = 0; q < P; q++) {
                auto it = pos.find((int64_t)ppl[ipp + q].y * nsp + ppl[ipp + q].x);
                if (it == pos.end()) return ccd::fail(CCD_E_INVALID, "pp channel %d lacks a swapped pair", k);
                pp_swap[ipp + q] = it->second;
            }
            pos.clear();
            for (int q = 0; q < H; q++) pos[(int64_t)hhl[ihh + q].x * nsp + hhl[ihh + q].y] = q;
            for (int q = 0; q < H; q++) {
                auto it = pos.find((int64_t)hhl[ihh + q].y * nsp + hhl[ihh + q].x);
                if (it == pos.end()) return ccd::fail(CCD_E_INVALID, "hh channel %d lacks a swapped pair", k);
                hh_swap[ihh + q] = it->second;
            }
            // packed rows: pairs a<b in list order
            const int pr_off = (int)prow.size();
            std::vector<int32_t> packed_of(P, -1);
            for (int q = 0; q < P; q++)
                if (ppl[ipp + q].x < ppl[ipp + q].y) {
                    packed_of[q] = (int32_t)(prow.size() - pr_off);
                    prow.push_back(q);
                    prow_sw.push_back(pp_swap[ipp + q]);
                }
            const int npk = (int)prow.size() - pr_off;
            for (int q = 0; q < P; q++) {
                const int a = ppl[ipp + q].x, b = ppl[ipp + q].y;
                if (a < b) pkc[ipp + q] = packed_of[q] << 1;
                else if (a > b) pkc[ipp + q] = (packed_of[pp_swap[ipp + q]] << 1) | 1;
            }
            const int ldv = (npk + 7) / 8 * 8;
            DChan ch{off_t, off_vp, off_tp, off_hhhh, P, H, (int32_t)ipp, (int32_t)ihh, npk, ldv, pr_off, 0};
            c->chans.push_back(ch);
            off_vp += (int64_t)npk * ldv;
            off_tp += (int64_t)npk * H;
            off_hhhh += (int64_t)H * H;
            c->n_vp_alg += (int64_t)npk * npk;
            c->flops += 2.0 * H * H * P + 2.0 * P * H * H + 2.0 * P * P * H;
            c->flops_pppp += 2.0 * P * P * H;
        }
        off_t += (int64_t)P * H;
        ipp += P; ihh += H;
    }
    c->nt2 = off_t;
    c->n_vp = off_vp; c->n_tp = off_tp; c->n_hhhh = off_hhhh;
    if (c->nt2 >= (int64_t)INT32_MAX) return ccd::fail(CCD_E_INVALID, "T2 has %lld elements (limit 2^31)", (long long)c->nt2);
    for (const DChan &ch : c->chans)
        if (ch.hh > NTMAX * 16) return ccd::fail(CCD_E_INVALID, "channel with hh=%d exceeds %d", ch.hh, NTMAX * 16);

    // mod channels and position tables
    std::vector<int32_t> rowbase((size_t)np * nh, -1), colidx((size_t)nh * np, -1);
    std::vector<int32_t> mod_of_ph((size_t)np * nh, -1), mod_of_hp((size_t)nh * np, -1);
    int64_t off_ph = 0, off_b = 0, iph = 0, ihp = 0;
    for (int k = 0; k < p->n_mod; k++) {
        const int PH = p->ph_dim[k], HP = p->hp_dim[k];
        for (int q = 0; q < PH; q++) {
            int a = phl[iph + q].x, i = phl[iph + q].y;
            size_t s = (size_t)(a - nh) * nh + i;
            if (mod_of_ph[s] >= 0) return ccd::fail(CCD_E_INVALID, "ph pair (%d,%d) listed twice", a, i);
            rowbase[s] = (int32_t)(off_ph + (int64_t)q * HP);
            mod_of_ph[s] = k;
        }
        for (int q = 0; q < HP; q++) {
            int j = hpl[ihp + q].x, b = hpl[ihp + q].y;
            size_t s = (size_t)j * np + (b - nh);
            if (mod_of_hp[s] >= 0) return ccd::fail(CCD_E_INVALID, "hp pair (%d,%d) listed twice", j, b);
            colidx[s] = q;
            mod_of_hp[s] = k;
        }
        if (PH > 0 && HP > 0) {
            DMod m{off_ph, off_b, PH, HP, (int32_t)iph, (int32_t)ihp};
            c->mods.push_back(m);
            off_b += (int64_t)HP * HP;
            c->flops += 2.0 * HP * HP * PH + 2.0 * PH * HP * HP;
        }
        off_ph += (int64_t)PH * HP;
        iph += PH; ihp += HP;
    }
    c->nph = off_ph; c->n_b = off_b;
    if (c->nph != c->nt2)
        return ccd::fail(CCD_E_INVALID, "phhp layout has %lld elements, pphh has %lld", (long long)c->nph,
                         (long long)c->nt2);
    // every T2 element must land in one mod block, (a,i) and (j,b) in the same channel
    for (const DChan &ch : c->chans)
        for (int AB = 0; AB < ch.pp; AB++)
            for (int IJ = 0; IJ < ch.hh; IJ++) {
                int a = ppl[ch.pp_off + AB].x, b = ppl[ch.pp_off + AB].y;
                int i = hhl[ch.hh_off + IJ].x, j = hhl[ch.hh_off + IJ].y;
                const int32_t m1 = mod_of_ph[(size_t)(a - nh) * nh + i], m2 = mod_of_hp[(size_t)j * np + (b - nh)];
                const int32_t m3 = mod_of_ph[(size_t)(b - nh) * nh + j], m4 = mod_of_hp[(size_t)i * np + (a - nh)];
                if (m1 < 0 || m1 != m2 || m3 < 0 || m3 != m4)
                    return ccd::fail(CCD_E_INVALID, "T2 element (%d,%d,%d,%d) has no consistent mod channel", a,
                                     b, i, j);
            }
    size_t lds = 0;
    for (const DMod &m : c->mods) lds = std::max(lds, (size_t)m.hp * m.hp * sizeof(double));
    if (lds > 64 * 1024) return ccd::fail(CCD_E_INVALID, "mod channel hp=%zu exceeds the ring kernel's LDS", lds);
    c->ring_lds_max = lds;

    // tiles: update tiles (~256 elements), ladder tiles (64 rows), X tiles (64 rows of hh)
    for (int k = 0; k < (int)c->chans.size(); k++) {
        const DChan &ch = c->chans[k];
        const int rows = std::max(1, kThreads / ch.hh);
        for (int r = 0; r < ch.pp; r += rows) c->utiles.push_back(DTile{k, r, std::min(rows, ch.pp - r), 0});
        for (int r = 0; r < ch.npk; r += LROWS) c->ltiles.push_back(DTile{k, r, std::min(LROWS, ch.npk - r), 0});
    }
    for (int k = 0; k < (int)c->chans.size(); k++)
        for (int r = 0; r < c->chans[k].hh; r += 16)
            for (int q = 0; q < c->chans[k].hh; q += 16) c->xtiles.push_back(DTile{k, r, 16, q});
    // one-body segments: chi_i over hh columns with first index i, psi_a over pp rows with first index a
    std::vector<std::vector<DSeg>> sg((size_t)nh + np);
    for (const DChan &ch : c->chans) {
        for (int IJ = 0; IJ < ch.hh; IJ++)
            sg[hhl[ch.hh_off + IJ].x].push_back(DSeg{ch.off_t + IJ, ch.hh, ch.pp});
        for (int AB = 0; AB < ch.pp; AB++)
            sg[(size_t)nh + (ppl[ch.pp_off + AB].x - nh)].push_back(DSeg{ch.off_t + (int64_t)AB * ch.hh, 1, ch.hh});
    }
    std::vector<int32_t> seg_ptr(1, 0);
    std::vector<DSeg> segs;
    for (auto &v : sg) {
        segs.insert(segs.end(), v.begin(), v.end());
        seg_ptr.push_back((int32_t)segs.size());
    }

    // --- device tables ---
    int rc;
    if ((rc = upload(c, &c->d_chans, c->chans)) || (rc = upload(c, &c->d_mods, c->mods)) ||
        (rc = upload(c, &c->d_ltiles, c->ltiles)) || (rc = upload(c, &c->d_utiles, c->utiles)) ||
        (rc = upload(c, &c->d_xtiles, c->xtiles)) ||
        (rc = upload(c, &c->d_ppl, ppl)) || (rc = upload(c, &c->d_hhl, hhl)) ||
        (rc = upload(c, &c->d_phl, phl)) || (rc = upload(c, &c->d_hpl, hpl)) ||
        (rc = upload(c, &c->d_pp_swap, pp_swap)) || (rc = upload(c, &c->d_hh_swap, hh_swap)) ||
        (rc = upload(c, &c->d_pkc, pkc)) || (rc = upload(c, &c->d_prow, prow)) ||
        (rc = upload(c, &c->d_prow_sw, prow_sw)) ||
        (rc = upload(c, &c->d_rowbase, rowbase)) || (rc = upload(c, &c->d_colidx, colidx)) ||
        (rc = upload(c, &c->d_seg_ptr, seg_ptr)) || (rc = upload(c, &c->d_segs, segs)))
        return rc;
    std::vector<double> eps(p->eps, p->eps + nsp);
    if ((rc = upload(c, &c->d_eps, eps))) return rc;
    if ((rc = dalloc(c, &c->d_onebody, (size_t)nsp))) return rc;
    for (int k = 0; k < 2; k++)
        if ((rc = dalloc(c, &c->d_t[k], (size_t)c->nt2))) return rc;
    if ((rc = dalloc(c, &c->d_d, (size_t)c->n_tp)) || (rc = dalloc(c, &c->d_tph, (size_t)c->nt2)) ||
        (rc = dalloc(c, &c->d_r, (size_t)c->nt2)) || (rc = dalloc(c, &c->d_w, (size_t)c->nt2)) ||
        (rc = dalloc(c, &c->d_x, (size_t)c->n_hhhh)) || (rc = dalloc(c, &c->d_vt, (size_t)c->nt2)) ||
        (rc = dalloc(c, &c->d_vh, (size_t)c->n_hhhh)) || (rc = dalloc(c, &c->d_vp, (size_t)c->n_vp)) ||
        (rc = dalloc(c, &c->d_vtp, (size_t)c->n_tp)) || (rc = dalloc(c, &c->d_tas, (size_t)c->n_tp)) ||
        (rc = dalloc(c, &c->d_va, (size_t)c->nt2)) || (rc = dalloc(c, &c->d_vb, (size_t)c->n_b)) ||
        (rc = dalloc(c, &c->d_partials, std::max(c->utiles.size(), (size_t)1))) ||
        (rc = dalloc(c, &c->d_scalar, 4)))
        return rc;
    HIP_TRY(hipHostMalloc((void **)&c->h_scalar, 4 * sizeof(double), hipHostMallocDefault));

    // --- interaction ---
    if (p->gen_kind != CCD_GEN_NONE) {
        std::vector<int32_t> qn(p->qn, p->qn + 5 * (size_t)nsp);
        if ((rc = upload(c, &c->d_qn, qn))) return rc;
        Gen G{p->gen_kind, p->gen_nmax, p->gen_g, c->d_qn, nullptr};
        if (p->gen_kind == CCD_GEN_ELECTRON_GAS) {
            // every momentum transfer of the basis must index the table
            for (int s = 0; s < nsp; s++)
                for (int k = 0; k < 3; k++)
                    if (std::abs(p->qn[5 * s + k]) > p->gen_nmax)
                        return ccd::fail(CCD_E_INVALID, "state %d has |n| > gen_nmax", s);
            const size_t w = 4 * (size_t)p->gen_nmax + 1;
            std::vector<double> tab(p->gen_table, p->gen_table + w * w * w);
            if ((rc = upload(c, &c->d_table, tab))) return rc;
            G.table = c->d_table;
        }
        if (!c->ltiles.empty())
            hipLaunchKernelGGL(k_gen_vp, dim3((unsigned)c->ltiles.size()), dim3(kThreads), 0, c->stream,
                               c->d_ltiles, c->d_chans, c->d_ppl, c->d_prow, G, c->d_vp);
        if (!c->chans.empty())
            hipLaunchKernelGGL(k_gen_direct_small, dim3((unsigned)c->chans.size()), dim3(kThreads), 0,
                               c->stream, c->d_chans, c->d_ppl, c->d_hhl, G, c->d_vt, c->d_vh);
        if (!c->mods.empty())
            hipLaunchKernelGGL(k_gen_mod, dim3((unsigned)c->mods.size()), dim3(kThreads), 0, c->stream,
                               c->d_mods, c->d_phl, c->d_hpl, G, c->d_va, c->d_vb);
        if ((rc = check_launch("generate V"))) return rc;
    } else {
        // caller V chains: copy the active blocks
        std::vector<int64_t> o_hhhh(p->n_chan + 1, 0), o_pppp(p->n_chan + 1, 0), o_b(p->n_mod + 1, 0);
        for (int k = 0; k < p->n_chan; k++) {
            o_hhhh[k + 1] = o_hhhh[k] + (int64_t)p->hh_dim[k] * p->hh_dim[k];
            o_pppp[k + 1] = o_pppp[k] + (int64_t)p->pp_dim[k] * p->pp_dim[k];
        }
        for (int k = 0; k < p->n_mod; k++) o_b[k + 1] = o_b[k] + (int64_t)p->hp_dim[k] * p->hp_dim[k];
        int ci = 0;
        std::vector<double> blk;
        for (int k = 0; k < p->n_chan; k++) {
            if (!(p->pp_dim[k] > 0 && p->hh_dim[k] > 0)) continue;
            const DChan &ch = c->chans[ci++];
            // The packed ladder needs V^pppp antisymmetric in both pairs (exactly,
            // as calcTBME produces it); check, then keep the a<b, c<d block.
            const double *V = p->v_pppp + o_pppp[k];
            const int P = ch.pp;
            for (int r = 0; r < P; r++) {
                const int rs = pp_swap[ch.pp_off + r];
                for (int q = 0; q < P; q++) {
                    const double v = V[(size_t)r * P + q];
                    if (V[(size_t)rs * P + q] != -v || V[(size_t)r * P + pp_swap[ch.pp_off + q]] != -v)
                        return ccd::fail(CCD_E_INVALID, "V^pppp of channel %d is not antisymmetric", k);
                }
            }
            blk.assign((size_t)ch.npk * ch.ldv, 0.0);
            for (int u = 0; u < ch.npk; u++)
                for (int w = 0; w < ch.npk; w++)
                    blk[(size_t)u * ch.ldv + w] = V[(size_t)prow[ch.pr_off + u] * P + prow[ch.pr_off + w]];
            HIP_TRY(hipMemcpy(c->d_vp + ch.off_vp, blk.data(), sizeof(double) * blk.size(), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(c->d_vh + ch.off_hhhh, p->v_hhhh + o_hhhh[k],
                              sizeof(double) * ch.hh * (size_t)ch.hh, hipMemcpyHostToDevice));
        }
        // V^hhpp has the pphh element count and offsets; stage through d_w, then transpose
        HIP_TRY(hipMemcpy(c->d_w, p->v_hhpp, sizeof(double) * (size_t)c->nt2, hipMemcpyHostToDevice));
        if (!c->chans.empty())
            hipLaunchKernelGGL(k_transpose_hhpp, dim3((unsigned)c->chans.size()), dim3(kThreads), 0,
                               c->stream, c->d_chans, c->d_w, c->d_vt);
        if ((rc = check_launch("k_transpose_hhpp"))) return rc;
        HIP_TRY(hipMemcpy(c->d_va, p->v_modA, sizeof(double) * (size_t)c->nt2, hipMemcpyHostToDevice));
        int mi = 0;
        for (int k = 0; k < p->n_mod; k++) {
            if (!(p->ph_dim[k] > 0 && p->hp_dim[k] > 0)) continue;
            const DMod &m = c->mods[mi++];
            HIP_TRY(hipMemcpy(c->d_vb + m.off_b, p->v_modB + o_b[k], sizeof(double) * m.hp * (size_t)m.hp,
                              hipMemcpyHostToDevice));
        }
    }
    if (!c->ltiles.empty()) {
        hipLaunchKernelGGL(k_pack_vt, dim3((unsigned)c->ltiles.size()), dim3(kThreads), 0, c->stream, c->d_ltiles,
                           c->d_chans, c->d_prow, c->d_vt, c->d_vtp);
        if ((rc = check_launch("k_pack_vt"))) return rc;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int zero_state(ccd_ctx *c) {
    for (int k = 0; k < 2; k++) HIP_TRY(hipMemsetAsync(c->d_t[k], 0, sizeof(double) * c->nt2, c->stream));
    HIP_TRY(hipMemsetAsync(c->d_tph, 0, sizeof(double) * c->nt2, c->stream));
    HIP_TRY(hipMemsetAsync(c->d_onebody, 0, sizeof(double) * c->nsp, c->stream));
    c->cur = 0;
    return CCD_OK;
}

// Load host amplitudes as the current T and derive T^ph, chi, psi from them.
int set_state(ccd_ctx *c, const double *t2) {
    HIP_TRY(hipMemcpyAsync(c->d_t[c->cur], t2, sizeof(double) * c->nt2, hipMemcpyHostToDevice, c->stream));
    UpdArgs a = upd_args(c, c->d_t[c->cur], nullptr, 0.0);
    if (!c->utiles.empty()) {
        hipLaunchKernelGGL(k_remap_ph, dim3((unsigned)c->utiles.size()), dim3(kThreads), 0, c->stream, a);
        int rc = check_launch("k_remap_ph");
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_hadamard, dim3(1024), dim3(kThreads), 0, c->stream, c->d_vt, c->d_t[c->cur], c->d_w,
                       c->nt2);
    int rc = check_launch("k_hadamard");
    if (rc) return rc;
    hipLaunchKernelGGL(k_onebody, dim3((unsigned)c->nsp), dim3(kThreads), 0, c->stream, c->d_seg_ptr,
                       c->d_segs, c->d_w, c->d_onebody);
    return check_launch("k_onebody");
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

int ccd_ctx_create(const ccd_problem *prob, int precision, ccd_ctx **out) {
    ccd::clear_error();
    if (!out) return ccd::fail(CCD_E_INVALID, "ccd_ctx_create: out is null");
    *out = nullptr;
    int rc = validate(prob);
    if (rc) return rc;
    if (precision != CCD_PREC_FP64)
        return ccd::fail(CCD_E_INVALID, "precision %d is not available", precision);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return ccd::fail(CCD_E_HIP, "no HIP device available");
    ccd_ctx *c = new (std::nothrow) ccd_ctx();
    if (!c) return ccd::fail(CCD_E_NOMEM, "out of host memory");
    c->precision = precision;
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ccd::fail(CCD_E_HIP, "could not create a HIP stream");
    }
    try {
        rc = build(c, prob);
    } catch (const std::bad_alloc &) {
        rc = ccd::fail(CCD_E_NOMEM, "out of host memory while building tables");
    }
    if (!rc) rc = zero_state(c);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = ccd::fail(CCD_E_HIP, "setup failed: %s", hipGetErrorString(e));
    }
    if (rc) {
        free_ctx(c);
        return rc;
    }
    *out = c;
    return CCD_OK;
}

int ccd_ctx_destroy(ccd_ctx *ctx) {
    free_ctx(ctx);
    return CCD_OK;
}

int ccd_ctx_info(const ccd_ctx *ctx, int64_t *n_t2, int64_t *device_bytes) {
    if (!ctx) return ccd::fail(CCD_E_INVALID, "null context");
    if (n_t2) *n_t2 = ctx->nt2;
    if (device_bytes) *device_bytes = ctx->bytes;
    return CCD_OK;
}

int ccd_work(const ccd_ctx *ctx, double *flops, double *flops_pppp, double *ladder_bytes) {
    if (!ctx) return ccd::fail(CCD_E_INVALID, "null context");
    if (flops) *flops = ctx->flops;
    if (flops_pppp) *flops_pppp = ctx->flops_pppp;
    if (ladder_bytes) {
        // one k_ladder_v launch reads the packed V^pppp, the antisymmetrised
        // amplitudes and X once and writes D_p once
        double b = 0.0;
        for (const DChan &ch : ctx->chans)
            b += (double)ch.npk * ch.npk + 2.0 * ch.npk * ch.hh + (double)ch.hh * ch.hh;
        *ladder_bytes = (ctx->precision == CCD_PREC_FP64 ? 8.0 : 4.0) * b;
    }
    return CCD_OK;
}

int ccd_mbpt2(ccd_ctx *c, double *corr) {
    ccd::clear_error();
    if (!c || !corr) return ccd::fail(CCD_E_INVALID, "ccd_mbpt2: null argument");
    if (c->utiles.empty()) {
        *corr = 0.0;
        return CCD_OK;
    }
    hipLaunchKernelGGL(k_mbpt2, dim3((unsigned)c->utiles.size()), dim3(kThreads), 0, c->stream, c->d_utiles,
                       c->d_chans, c->d_ppl, c->d_hhl, c->d_eps, c->d_vt, c->d_partials);
    int rc = check_launch("k_mbpt2");
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, c->stream, c->d_partials,
                       (int)c->utiles.size(), 0.25, c->d_scalar);
    if ((rc = check_launch("k_sum_partials"))) return rc;
    return read_scalar(c, corr);
}

int ccd_reset(ccd_ctx *c) {
    ccd::clear_error();
    if (!c) return ccd::fail(CCD_E_INVALID, "null context");
    int rc = zero_state(c);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int ccd_iterate(ccd_ctx *c, double mixing, double *corr) {
    ccd::clear_error();
    if (!c) return ccd::fail(CCD_E_INVALID, "null context");
    int rc = run_iteration(c, mixing, nullptr);
    if (rc) return rc;
    double e = 0.0;
    if ((rc = read_scalar(c, &e))) return rc;
    if (corr) *corr = e;
    return CCD_OK;
}

int ccd_solve(ccd_ctx *c, double tol, int max_iters, double mixing, double *energies, int cap, int *n_iter,
              int *stop) {
    ccd::clear_error();
    if (!c) return ccd::fail(CCD_E_INVALID, "null context");
    int rc = zero_state(c);
    if (rc) return rc;
    double diff = 1.0, prev = 100.0, corr = 100.0;
    int it = 0, why = CCD_STOP_CONVERGED;
    while (std::fabs(diff) > tol) {
        if (it > max_iters) {
            why = CCD_STOP_MAX_ITERS;
            break;
        }
        if ((rc = run_iteration(c, mixing, nullptr))) return rc;
        if ((rc = read_scalar(c, &corr))) return rc;
        diff = corr - prev;
        prev = corr;
        if (energies && it < cap) energies[it] = corr;
        it++;
    }
    if (why == CCD_STOP_CONVERGED && std::isnan(diff)) why = CCD_STOP_NAN;
    if (n_iter) *n_iter = it;
    if (stop) *stop = why;
    return CCD_OK;
}

int ccd_get_t2(ccd_ctx *c, double *t2) {
    ccd::clear_error();
    if (!c || !t2) return ccd::fail(CCD_E_INVALID, "ccd_get_t2: null argument");
    HIP_TRY(hipMemcpyAsync(t2, c->d_t[c->cur], sizeof(double) * c->nt2, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int ccd_set_t2(ccd_ctx *c, const double *t2) {
    ccd::clear_error();
    if (!c || !t2) return ccd::fail(CCD_E_INVALID, "ccd_set_t2: null argument");
    int rc = set_state(c, t2);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int ccd_ladder(ccd_ctx *c, const double *t, double *d) {
    ccd::clear_error();
    if (!c || !t || !d) return ccd::fail(CCD_E_INVALID, "ccd_ladder: null argument");
    double *dt = c->d_t[1 - c->cur];  // scratch: the non-current T buffer
    HIP_TRY(hipMemcpyAsync(dt, t, sizeof(double) * c->nt2, hipMemcpyHostToDevice, c->stream));
    int rc;
    if ((rc = launch_ladder_x(c, dt)) || (rc = launch_ladder_d(c, dt))) return rc;
    // expand D_p to the ordered pphh layout through the sign codes
    UpdArgs a = upd_args(c, nullptr, c->d_w, 0.0);
    hipLaunchKernelGGL(k_expand_d, dim3((unsigned)c->utiles.size()), dim3(kThreads), 0, c->stream, a);
    if ((rc = check_launch("k_expand_d"))) return rc;
    HIP_TRY(hipMemcpyAsync(d, c->d_w, sizeof(double) * c->nt2, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int ccd_ring(ccd_ctx *c, const double *tph, double *r) {
    ccd::clear_error();
    if (!c || !tph || !r) return ccd::fail(CCD_E_INVALID, "ccd_ring: null argument");
    HIP_TRY(hipMemcpyAsync(c->d_tph, tph, sizeof(double) * c->nt2, hipMemcpyHostToDevice, c->stream));
    int rc = launch_ring(c);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(r, c->d_r, sizeof(double) * c->nt2, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int ccd_remap_ph(ccd_ctx *c, const double *t, double *tph) {
    ccd::clear_error();
    if (!c || !t || !tph) return ccd::fail(CCD_E_INVALID, "ccd_remap_ph: null argument");
    double *dt = c->d_t[1 - c->cur];
    HIP_TRY(hipMemcpyAsync(dt, t, sizeof(double) * c->nt2, hipMemcpyHostToDevice, c->stream));
    UpdArgs a = upd_args(c, dt, nullptr, 0.0);
    if (!c->utiles.empty()) {
        hipLaunchKernelGGL(k_remap_ph, dim3((unsigned)c->utiles.size()), dim3(kThreads), 0, c->stream, a);
        int rc = check_launch("k_remap_ph");
        if (rc) return rc;
    }
    HIP_TRY(hipMemcpyAsync(tph, c->d_tph, sizeof(double) * c->nt2, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int ccd_onebody(ccd_ctx *c, const double *t, double *chi, double *psi) {
    ccd::clear_error();
    if (!c || !t || !chi || !psi) return ccd::fail(CCD_E_INVALID, "ccd_onebody: null argument");
    int rc = set_state(c, t);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(chi, c->d_onebody, sizeof(double) * c->nh, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(psi, c->d_onebody + c->nh, sizeof(double) * c->np, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int ccd_step(ccd_ctx *c, const double *t_old, double *t_new, double mixing, double *corr) {
    ccd::clear_error();
    if (!c || !t_old || !t_new) return ccd::fail(CCD_E_INVALID, "ccd_step: null argument");
    int rc = set_state(c, t_old);
    if (rc) return rc;
    if ((rc = run_iteration(c, mixing, nullptr))) return rc;
    double e = 0.0;
    if ((rc = read_scalar(c, &e))) return rc;
    HIP_TRY(hipMemcpyAsync(t_new, c->d_t[c->cur], sizeof(double) * c->nt2, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (corr) *corr = e;
    return CCD_OK;
}

int ccd_time_phases(ccd_ctx *c, int iters, double mixing, double *ms) {
    ccd::clear_error();
    if (!c || !ms || iters <= 0) return ccd::fail(CCD_E_INVALID, "ccd_time_phases: bad argument");
    hipEvent_t ev[6];
    for (int k = 0; k < 6; k++) HIP_TRY(hipEventCreate(&ev[k]));
    double acc[6] = {0, 0, 0, 0, 0, 0};
    int rc = CCD_OK;
    for (int it = 0; it < iters && !rc; it++) {
        rc = run_iteration(c, mixing, ev);
        if (rc) break;
        hipError_t e = hipEventSynchronize(ev[5]);
        if (e != hipSuccess) {
            rc = ccd::fail(CCD_E_HIP, "iteration failed: %s", hipGetErrorString(e));
            break;
        }
        for (int k = 0; k < 5; k++) {
            float f = 0.f;
            (void)hipEventElapsedTime(&f, ev[k], ev[k + 1]);
            acc[k] += f;
        }
        float f = 0.f;
        (void)hipEventElapsedTime(&f, ev[0], ev[5]);
        acc[5] += f;
    }
    for (int k = 0; k < 6; k++) (void)hipEventDestroy(ev[k]);
    if (rc) return rc;
    for (int k = 0; k < 6; k++) ms[k] = acc[k] / iters;
    return CCD_OK;
}

}  // extern "C"
