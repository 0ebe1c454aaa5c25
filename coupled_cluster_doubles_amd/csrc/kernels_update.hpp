// Amplitude update, remaps and reductions:
//   calcChannelT2_pphh (src/calc.cpp:368-421)              -> k_update
//   the one-body terms (calc.cpp:9-25, 73-89)              -> chi/psi from k_reduce, applied in k_update
//   mapToDistributed(ai_jb) / mapFromDistributed x 4       -> k_transpose (per particle, LDS-staged)
//   (src/Chain.hpp:710-802)
//   the energy sum (calc.cpp:417-420, main.cpp:667-668)     -> k_reduce (last block)
//   MBPT2ChannelCorr (src/MBPT2Corr.cpp:6-28)              -> k_mbpt2
#pragma once

#include "device_common.hpp"

// Row pairs: the first npk entries of a channel's pair list are the packed rows
// u (pair a<b, partner row (b,a)); the next ndiag entries are rows (a,a),
// paired with themselves.  An update tile is a run of pairs of one channel
// (row0 = first pair, nrows = number of pairs, aux = offset of its column
// partial sums).
struct UpdArgs {
    const DTile *tiles;
    const DChan *chans;
    const int2 *ppl, *hhl;
    const int32_t *prow, *prow_sw;   // row of (a,b) and of (b,a) per pair
    const double *eps, *chi, *psi;   // chi [nh], psi [np]
    const double *vt;                // <ij||ab> in the pphh layout
    const float *vt32;               // its FP32 rounding (what the update reads in FP32 iterations)
    const void *dp;                  // ladder D on packed rows: one slab per k segment, then 1/2 T X
    const float *dsym;               // FP32 iterations with class 1 on LK_BF16X3SYM: its symmetric
                                     // partial sums (rows of row tile J: J slabs), else nullptr
    const void *g;                   // ring R gathered to the pphh layout: g[(a,b),(i,j)] = R[(a,i),(j,b)]
                                     // (both in the contractions' arithmetic type)
    const double *t_old;
    double *t_new;                   // new amplitudes
    void *tas;                       // their (a,b)-antisymmetric packed rows, in OT
    double *red;                     // W = V o T partial sums, each tile's stored contiguously:
                                     //   the column partials at [tile aux + IJ], then,
    int64_t n_cp;                    //   from n_cp on, the row sums at [pp_off + row]
    double mixing;
    int nh;
    const LoopCtl *ctl;              // loop state of the context
    const LoopCtl *pctl;             // per solve (a batch); pctl == ctl for a single solve
};

// The symmetric partial sums of the pppp ladder for packed row pr, column
// IJ of a class-1 channel (the kernel kind LK_BF16X3SYM): the rows of row
// tile J (64 rows) hold J slabs, those of the tiles before it, added in
// order (loads issued four at a time; the clamped loads are not added).
__device__ __forceinline__ double sym_rows(const float *__restrict__ dsym, const DChan &ch, int pr, int IJ) {
    const int H = ch.hh, J = pr >> 6;
    const float *s = dsym + (int64_t)ch.sym * 1024 + ((int64_t)(J * (J - 1) / 2) * 64 + (pr - 64 * J)) * H + IJ;
    const int64_t slab = (int64_t)64 * H;
    double d = 0.0;
    for (int j0 = 0; j0 < J; j0 += 4) {
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = s[min(j0 + u, J - 1) * slab];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (j0 + u < J) d += (double)v[u];
    }
    return d;
}

// UPDATE = true: one Jacobi step for both rows of each pair; UPDATE = false:
// only the W sums of the current amplitudes (to derive chi, psi, E from them),
// with V in the same arithmetic as the iterations that follow.
// A solve of a batch that has stopped keeps its amplitudes: its new T is the
// old one (written while the other buffer may still hold older amplitudes).  OT: type of the ladder and ring outputs (float after FP32
// contractions, which they hold exactly); the update itself runs in FP64.
template <bool UPDATE, typename OT>
__global__ __launch_bounds__(kThreads) void k_update(UpdArgs a) {
    const DTile tl = a.tiles[blockIdx.x];
    if (UPDATE && a.ctl->done) return;   // the W sums alone run whatever the loop state
    const OT *gg = static_cast<const OT *>(a.g);
    __shared__ double w1s[kThreads], w2s[kThreads], cs[kThreads];
    const DChan ch = a.chans[tl.chan];
    const bool frozen = UPDATE && a.pctl[ch.prob].done;
    // A solve of a batch that stopped two or more iterations ago: both
    // amplitude buffers and the packed rows already hold its final
    // amplitudes, and nothing reads its W sums until it resumes (which
    // rebuilds chi and psi), so its tiles have nothing to do.  (The reductions
    // that update the flags have finished before the update starts.)
    if (frozen && a.pctl[ch.prob].idle >= 2) return;
    const int H = ch.hh, n = tl.nrows * H;
    const int l = threadIdx.x & (kThreads - 1);
    // places of this slot's sums (a tile's column partials and row sums are
    // contiguous, so its stores coalesce; the reductions gather them)
    int rp1 = 0, rp2 = 0, cp = 0;
    bool pdiag = true;
    if (l < n) {
        // every load of the element pair is issued before the first store (the
        // stores alias nothing the loads read, but the compiler cannot know
        // that)
        const int pr = tl.row0 + l / H, IJ = l % H;
        const int JI = IJ < ch.hh2 ? (IJ ^ 1) : IJ;   // column (j,i) of column (i,j)
        // rows of the pair in the internal order: packed pair u at 2u (a,b) and
        // 2u+1 (b,a), the pairs (a,a) after them (the prow / prow_sw lists)
        const bool diag = pr >= ch.npk;
        const int r1 = diag ? ch.npk + pr : 2 * pr, r2 = diag ? r1 : r1 + 1;
        cp = tl.aux + IJ;
        rp1 = ch.pp_off + r1;
        rp2 = ch.pp_off + r2;
        pdiag = diag;
        const int64_t b1 = ch.off_t + (int64_t)r1 * H, b2 = ch.off_t + (int64_t)r2 * H;
        // V in the arithmetic of the iteration: FP32 iterations add its FP32
        // rounding, as their contractions are FP32 (the sums stay FP64)
        double v1, v2;
        if constexpr (sizeof(OT) == 4) {
            v1 = a.vt32[b1 + IJ];
            v2 = a.vt32[b2 + IJ];
        } else {
            v1 = a.vt[b1 + IJ];
            v2 = a.vt[b2 + IJ];
        }
        const double t1 = a.t_old[b1 + IJ], t2 = a.t_old[b2 + IJ];
        double w1, w2;
        if (UPDATE) {
            const int2 ab = a.ppl[ch.pp_off + r1], ij = a.hhl[ch.hh_off + IJ];
            const int ia = ab.x, ib = ab.y, ii = ij.x, jj = ij.y;
            const double t1s = a.t_old[b1 + JI], t2s = a.t_old[b2 + JI];
            const double g1 = gg[b1 + IJ], g1s = gg[b1 + JI], g2 = gg[b2 + IJ], g2s = gg[b2 + JI];
            // D_p: the slabs 0..nseg added in order, loaded four at a time
            // (slab indices clamped to nseg; the clamped loads are not added)
            double d = 0.0;
            if (!diag) {
                const int64_t slab = (int64_t)ch.npk * H;
                const OT *dq = static_cast<const OT *>(a.dp) + ch.off_dp + (int64_t)pr * H + IJ;
                for (int q0 = 0; q0 <= ch.nseg; q0 += 4) {
                    OT dv[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) dv[u] = dq[min(q0 + u, ch.nseg) * slab];
#pragma unroll
                    for (int u = 0; u < 4; u++)
                        if (q0 + u <= ch.nseg) d += (double)dv[u];
                }
                if (a.dsym && ch.sym >= 0) d += sym_rows(a.dsym, ch, pr, IJ);
            }
            const double ci = 0.5 * a.chi[ii], cj = 0.5 * a.chi[jj];
            const double pa = -0.5 * a.psi[ia - a.nh], pb = -0.5 * a.psi[ib - a.nh];
            const double ei = a.eps[ii], ej = a.eps[jj], ea = a.eps[ia], eb = a.eps[ib];
            // element (a,b,i,j), in the order of src/calc.cpp:393-408
            double x1 = v1;
            x1 += d;
            x1 += g1;
            x1 -= g1s;
            x1 -= g2;
            x1 += g2s;
            x1 += ci * t1s;
            x1 -= cj * t1;
            x1 += pb * t1;
            x1 -= pa * t2;
            x1 = x1 / (ei + ej - ea - eb);
            x1 = a.mixing * x1 + (1. - a.mixing) * t1;
            if (frozen) x1 = t1;
            w1 = v1 * x1;
            w2 = 0.0;
            double x2 = 0.0;
            if (!diag) {
                // element (b,a,i,j)
                x2 = v2;
                x2 -= d;
                x2 += g2;
                x2 -= g2s;
                x2 -= g1;
                x2 += g1s;
                x2 += ci * t2s;
                x2 -= cj * t2;
                x2 += pa * t2;
                x2 -= pb * t1;
                x2 = x2 / (ei + ej - eb - ea);
                x2 = a.mixing * x2 + (1. - a.mixing) * t2;
                if (frozen) x2 = t2;
                w2 = v2 * x2;
            }
            a.t_new[b1 + IJ] = x1;
            if (!diag) {
                a.t_new[b2 + IJ] = x2;
                static_cast<OT *>(a.tas)[ch.off_tp + (int64_t)pr * H + IJ] = (OT)(0.5 * (x1 - x2));
            }
        } else {
            w1 = v1 * t1;
            w2 = diag ? 0.0 : v2 * t2;
        }
        w1s[l] = w1;
        w2s[l] = w2;
        cs[l] = w1 + w2;
    }
    __syncthreads();
    // deterministic pairwise trees: row sums of w1 and w2 over the columns of
    // each pair (slot IJ = 0 ends with the sum) and column partials of
    // w1 + w2 over the pairs of the tile (slot pr = 0); in step d a slot adds
    // the one d further on, which does not change in that step
    const int pq = l / H, cq = l - pq * H;
    for (int d = 1; d < H || d < tl.nrows; d <<= 1) {
        if (l < n) {
            if ((cq & (2 * d - 1)) == 0 && cq + d < H) {
                w1s[l] += w1s[l + d];
                w2s[l] += w2s[l + d];
            }
            if ((pq & (2 * d - 1)) == 0 && pq + d < tl.nrows) cs[l] += cs[l + d * H];
        }
        __syncthreads();
    }
    if (l < n && cq == 0) {
        a.red[a.n_cp + rp1] = w1s[l];
        if (!pdiag) a.red[a.n_cp + rp2] = w2s[l];
    }
    if (l < n && pq == 0) a.red[cp] = cs[l];
}

// Per-particle transposition between the pphh and phhp layouts (both hold
// every amplitude once; the block of particle a holds the elements whose
// first particle index is a, base[a] .. base[a+1]).  Slot l of the source
// order is loaded from src[src_addr[base+l]] into LDS (row runs, so the loads
// coalesce), then slot l of the destination order is written to
// dst[dst_addr[base+l]] from LDS slot perm[base+l].  With src = pphh it is the
// reference's mapToDistributed(ai_jb); with src = phhp it gathers R into the
// pphh layout.  Elements are converted from S to E on the way (T^ph for FP32
// contractions is the FP32 rounding of T, which is all the FP32 ring reads).
// The LDS block holds the largest particle block plus one spare slot.  In a
// batch (pctl, pprob: the solve of each particle) the blocks of stopped
// solves return at entry.
template <typename S, typename E = S>
__global__ __launch_bounds__(kThreads) void k_transpose(const int32_t *__restrict__ base,
                                                        const int32_t *__restrict__ src_addr,
                                                        const int32_t *__restrict__ dst_addr,
                                                        const uint16_t *__restrict__ perm,
                                                        const S *__restrict__ src,
                                                        E *__restrict__ dst,
                                                        const LoopCtl *__restrict__ ctl,
                                                        const int32_t *__restrict__ pprob,
                                                        const LoopCtl *__restrict__ pctl) {
    if (ctl->done) return;
    extern __shared__ double tr_lds[];
    E *buf = reinterpret_cast<E *>(tr_lds);
    const int a = blockIdx.x;
    if (pctl && block_solve_stopped(pctl, pprob[a])) return;   // particle a's solve (a batch) has stopped
    const int b0 = base[a], n = base[a + 1] - b0;
    // rounds of U slots per thread: every address load of a round is issued,
    // then every data load, so a thread waits for two memory round trips per
    // round instead of two per slot (a masked-off slot reads element 0 of its
    // array and discards it)
    constexpr int U = 8;
    for (int l0 = threadIdx.x; l0 < n; l0 += kThreads * U) {
        int ad[U];
        E v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int l = l0 + u * kThreads;
            ad[u] = src_addr[b0 + (l < n ? l : 0)];
        }
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = (E)src[ad[u]];
        // slots past the block go to the spare LDS slot n (never read), so
        // that no load sits behind a branch
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int l = l0 + u * kThreads;
            buf[l < n ? l : n] = v[u];
        }
    }
    __syncthreads();
    for (int l0 = threadIdx.x; l0 < n; l0 += kThreads * U) {
        int ad[U], pm[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int l = l0 + u * kThreads, q = b0 + (l < n ? l : 0);
            ad[u] = dst_addr[q];
            pm[u] = perm[q];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int l = l0 + u * kThreads;
            if (l < n) dst[ad[u]] = buf[pm[u]];
        }
    }
}

// chi_i (blocks 0..nh-1) and psi_a (blocks nh..nsp-1) as fixed-order sums of
// their column partials / row sums: parts[idx[q]] for q in ptr[k] ..
// ptr[k+1] (the update stores each tile's sums contiguously, the lists
// gather them).  The chi block that finishes last
// (ticket counter) then forms, for every solve p still running, E_p = 1/4
// sum of chi_i over its holes [hlo[p], hlo[p+1]) (= 1/4 sum V o T: every
// element has one first hole index) in a fixed order, and runs the loop
// bookkeeping of src/main.cpp:511-519, 667-670 on pctl[p]: dE = E - E_prev,
// E_prev = E, iterations++, and the loop goes on while |dE| > tol unless
// more than max_iters iterations have run.  A single solve has pctl == ctl;
// a batch counts its iterations in ctl and stops once every solve has.  The
// last block resets the ticket counter for the next launch.  Without holes
// (an empty problem) block 0 does that part.  bookkeep = 0: chi and psi only
// (rebuilt from the current amplitudes), whatever the loop state, no energy
// and no loop step.
__global__ __launch_bounds__(kThreads) void k_reduce(const int32_t *__restrict__ ptr,
                                                     const int32_t *__restrict__ idx, int nh,
                                                     const double *__restrict__ parts,
                                                     double *__restrict__ out, int *__restrict__ ticket,
                                                     double *__restrict__ energy, LoopCtl *ctl,
                                                     double *__restrict__ log, int nprob,
                                                     const int32_t *__restrict__ hlo, LoopCtl *pctl,
                                                     int bookkeep) {
    if (bookkeep && ctl->done) return;
    __shared__ double red[kThreads / 64];
    __shared__ int last, pdone;
    const int k = blockIdx.x;
    // thread t adds list entries lo + t, lo + t + 256, ... in order; the
    // index and then the value loads of 8 strides are issued before their
    // adds (a slot past the list reads entry lo and is not added), so a
    // thread waits for two memory round trips per 8 elements
    constexpr int U = 8;
    const int lo = ptr[k], hi = ptr[k + 1];
    double s = 0.0;
    for (int q0 = lo + threadIdx.x; q0 < hi; q0 += kThreads * U) {
        int ix[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int q = q0 + u * kThreads;
            ix[u] = idx[q < hi ? q : lo];
        }
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = parts[ix[u]];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (q0 + u * kThreads < hi) s += v[u];
    }
    const double b = block_sum(s, red);
    if (threadIdx.x == 0) {
        out[k] = b;
        if (k < nh) {           // only the chi blocks take a ticket
            __threadfence();
            last = atomicAdd(ticket, 1) == nh - 1;
        } else {
            last = nh == 0 && k == 0;   // no holes (no amplitudes): block 0 runs the loop step, E = 0
        }
    }
    __syncthreads();
    if (!last) return;
    if (!bookkeep) {
        if (threadIdx.x == 0) *ticket = 0;
        return;
    }
    __threadfence();
    for (int p = 0; p < nprob; p++) {
        if (threadIdx.x == 0) pdone = pctl[p].done;
        __syncthreads();
        const int skip = pdone;
        __syncthreads();
        if (skip) {
            if (threadIdx.x == 0) pctl[p].idle += 1;
            continue;
        }
        double c = 0.0;
        for (int i = hlo[p] + threadIdx.x; i < hlo[p + 1]; i += kThreads)
            c += __hip_atomic_load(out + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const double tot = block_sum(c, red);
        if (threadIdx.x == 0) {
            LoopCtl *pc = pctl + p;
            const double e = 0.25 * tot;
            energy[p] = e;
            const double diff = e - pc->prev;
            pc->prev = e;
            const int it = pc->count;
            if (it < pc->cap) log[(int64_t)p * pc->cap + it] = e;
            pc->count = it + 1;
            if (!(fabs(diff) > pc->tol)) {
                pc->reason = isnan(diff) ? CCD_STOP_NAN : CCD_STOP_CONVERGED;
                pc->done = 1;
                pc->idle = 0;
            } else if (it + 1 > pc->max_iters) {
                pc->reason = CCD_STOP_MAX_ITERS;
                pc->done = 1;
                pc->idle = 0;
            }
        }
    }
    if (threadIdx.x == 0) {
        *ticket = 0;
        if (pctl != ctl) {
            int all = 1;
            for (int p = 0; p < nprob; p++) all &= pctl[p].done;
            ctl->count += 1;
            if (all) ctl->done = 1;
        }
    }
}

// MBPT2 terms v^2 / denominator over the pair tiles, per-block partial sums.
__global__ __launch_bounds__(kThreads) void k_mbpt2(UpdArgs a, double *__restrict__ partials) {
    __shared__ double red[kThreads / 64];
    const DTile tl = a.tiles[blockIdx.x];
    const DChan ch = a.chans[tl.chan];
    const int H = ch.hh, l = threadIdx.x;
    double s = 0.0;
    if (l < tl.nrows * H) {
        const int pr = tl.row0 + l / H, IJ = l % H;
        const int r1 = a.prow[ch.pr_off + pr], r2 = a.prow_sw[ch.pr_off + pr];
        const int2 ab = a.ppl[ch.pp_off + r1], ij = a.hhl[ch.hh_off + IJ];
        const double v1 = a.vt[ch.off_t + (int64_t)r1 * H + IJ];
        s = v1 * v1 / (a.eps[ij.x] + a.eps[ij.y] - a.eps[ab.x] - a.eps[ab.y]);
        if (r1 != r2) {
            const double v2 = a.vt[ch.off_t + (int64_t)r2 * H + IJ];
            s += v2 * v2 / (a.eps[ij.x] + a.eps[ij.y] - a.eps[ab.y] - a.eps[ab.x]);
        }
    }
    const double b = block_sum(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

// out = scale * sum(partials[0..n)) in a fixed order (one block).
__global__ __launch_bounds__(kThreads) void k_sum_partials(const double *__restrict__ partials, int n,
                                                           double scale, double *__restrict__ out) {
    __shared__ double red[kThreads / 64];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += kThreads) s += partials[k];
    const double b = block_sum(s, red);
    if (threadIdx.x == 0) out[0] = scale * b;
}

// Antisymmetrised packed rows of amplitudes: tas[u] = (t[(a,b)] - t[(b,a)]) / 2,
// stored in TT (the arithmetic type of the ladder that reads them).
template <typename TT>
__global__ __launch_bounds__(kThreads) void k_antisym(const DTile *__restrict__ tiles,
                                                      const DChan *__restrict__ chans,
                                                      const int32_t *__restrict__ prow,
                                                      const int32_t *__restrict__ prow_sw,
                                                      const double *__restrict__ t,
                                                      TT *__restrict__ tas) {
    const DTile tl = tiles[blockIdx.x];
    const DChan ch = chans[tl.chan];
    const int H = ch.hh;
    for (int l = threadIdx.x; l < tl.nrows * H; l += kThreads) {
        const int u = tl.row0 + l / H, ij = l % H;
        const double x = t[ch.off_t + (int64_t)prow[ch.pr_off + u] * H + ij];
        const double y = t[ch.off_t + (int64_t)prow_sw[ch.pr_off + u] * H + ij];
        tas[ch.off_tp + (int64_t)u * H + ij] = (TT)(0.5 * (x - y));
    }
}

// D in the ordered pphh layout from the packed ladder output (ladder entry point).
template <typename OT>
__global__ __launch_bounds__(kThreads) void k_expand_d(UpdArgs a, double *__restrict__ d) {
    const DTile tl = a.tiles[blockIdx.x];
    const DChan ch = a.chans[tl.chan];
    const int H = ch.hh, l = threadIdx.x;
    if (l >= tl.nrows * H) return;
    const int pr = tl.row0 + l / H, IJ = l % H;
    const int r1 = a.prow[ch.pr_off + pr], r2 = a.prow_sw[ch.pr_off + pr];
    if (r1 == r2) {
        d[ch.off_t + (int64_t)r1 * H + IJ] = 0.0;
    } else {
        const int64_t slab = (int64_t)ch.npk * H;
        const OT *dq = static_cast<const OT *>(a.dp) + ch.off_dp + (int64_t)pr * H + IJ;
        double v = 0.0;
        for (int q = 0; q <= ch.nseg; q++) v += (double)dq[q * slab];
        if (a.dsym && ch.sym >= 0) v += sym_rows(a.dsym, ch, pr, IJ);
        d[ch.off_t + (int64_t)r1 * H + IJ] = v;
        d[ch.off_t + (int64_t)r2 * H + IJ] = -v;
    }
}

// X of the channels of more than one k segment: V^hhhh plus the segments'
// partial blocks in segment order, one 16x16 tile per block (tile.row0,
// tile.aux: its first row and column).
__global__ __launch_bounds__(kThreads) void k_fold_x(const DTile *__restrict__ tiles,
                                                     const DChan *__restrict__ chans,
                                                     const double *__restrict__ vh,
                                                     const double *__restrict__ xp, double *__restrict__ x,
                                                     const LoopCtl *__restrict__ ctl,
                                                     const LoopCtl *__restrict__ pctl) {
    if (ctl->done) return;
    const DTile tl = tiles[blockIdx.x];
    const DChan ch = chans[tl.chan];
    if (wave_solve_stopped(pctl, ch.prob)) return;   // no barrier below: a per-wave answer suffices
    const int H = ch.hh, row = tl.row0 + (threadIdx.x >> 4), col = tl.aux + (threadIdx.x & 15);
    if (row >= H || col >= H) return;
    const int64_t o = (int64_t)row * H + col, hh2 = (int64_t)H * H;
    // the segments' loads issued four at a time (a clamped load is not added)
    double v = vh[ch.off_hhhh + o];
    for (int q0 = 0; q0 < ch.nxs; q0 += 4) {
        double t[4];
#pragma unroll
        for (int u = 0; u < 4; u++) t[u] = xp[ch.off_xp + min(q0 + u, ch.nxs - 1) * hh2 + o];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (q0 + u < ch.nxs) v += t[u];
    }
    x[ch.off_hhhh + o] = v;
}
