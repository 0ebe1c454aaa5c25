// Setup-time kernels: the interaction blocks (generated on the device or
// rearranged from caller-supplied chains).  They replace the reference's
// loadVnn_* loops (src/load.cpp:8-318) for the blocks the iteration reads.
#pragma once

#include "device_common.hpp"

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

// The FP32 packed V^pppp of the FP32 iterations with a stored V: the element
// the generated FP32 ladders build (gen_tbme_f32), so the two modes agree.
__global__ __launch_bounds__(kThreads) void k_gen_vp32(const DTile *__restrict__ tiles,
                                                       const DChan *__restrict__ chans,
                                                       const int2 *__restrict__ ppl,
                                                       const int32_t *__restrict__ prow, Gen G,
                                                       float *__restrict__ vp) {
    const DTile tl = tiles[blockIdx.x];
    const DChan ch = chans[tl.chan];
    const int ldv = ch.ldv, npk = ch.npk;
    const int64_t n = (int64_t)tl.nrows * ldv;
    for (int64_t l = threadIdx.x; l < n; l += kThreads) {
        const int u = tl.row0 + (int)(l / ldv), w = (int)(l % ldv);
        float v = 0.f;
        if (w < npk) {
            const int2 ab = ppl[ch.pp_off + prow[ch.pr_off + u]];
            const int2 cd = ppl[ch.pp_off + prow[ch.pr_off + w]];
            v = gen_tbme_f32(G, ab.x, ab.y, cd.x, cd.y);
        }
        vp[ch.off_vp + (int64_t)u * ldv + w] = v;
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

// Caller V^hhpp (hh x pp) -> pphh layout (internal rows and columns).
__global__ __launch_bounds__(kThreads) void k_transpose_hhpp(const DChan *__restrict__ chans,
                                                             const int32_t *__restrict__ pp_old,
                                                             const int32_t *__restrict__ hh_old,
                                                             const double *__restrict__ vhhpp,
                                                             double *__restrict__ vt) {
    const DChan ch = chans[blockIdx.x];
    const int P = ch.pp, H = ch.hh;
    for (int l = threadIdx.x; l < P * H; l += kThreads) {
        // internal row AB is the caller's row pp_old[AB], column IJ its column hh_old[IJ]
        const int AB = l / H, IJ = l - AB * H;
        vt[ch.off_t + l] = vhhpp[ch.off_t + (int64_t)hh_old[ch.hh_off + IJ] * P + pp_old[ch.pp_off + AB]];
    }
}

// dst[i] = src[idx[i]]: amplitudes between the caller's and the internal row order.
__global__ __launch_bounds__(kThreads) void k_gather(const double *__restrict__ src, const int32_t *__restrict__ idx,
                                                     double *__restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        dst[i] = src[idx[i]];
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

// FP32 copy of an FP64 array (round to nearest), for the mixed-precision
// contractions.
__global__ __launch_bounds__(kThreads) void k_to_f32(const double *__restrict__ src, float *__restrict__ dst,
                                                     int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        dst[i] = (float)src[i];
}
