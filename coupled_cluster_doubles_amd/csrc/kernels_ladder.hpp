// Particle-particle / hole-hole ladders and the hhhh quadratic term
// (calcChannelT2Diff_pphh, src/calc.cpp:132-203).
//
// With t and V antisymmetric in (a,b) and (c,d), the reference's ordered-pair
// ladder 1/2 sum_{cd} V[ab][cd] t[cd][ij] equals sum_{c<d} V[ab][cd] t[cd][ij],
// and D[(b,a)] = -D[(a,b)], D[(a,a)] = 0.  Both ladder kernels therefore run on
// the packed pairs a<b: a quarter of the V^pppp bytes and flops.  The
// contractions use v_mfma_f64_16x16x4f64 (layout in device_common.hpp).
#pragma once

#include <type_traits>

#include "device_common.hpp"

constexpr int KC = 32;                 // k per chunk of k_ladder_x (k-slot q of step s: k = 8q + s)
constexpr int KXSEG = 256;             // packed pairs per k segment of X (a partial X block each)
constexpr int NTMAX = 8;               // 16-column tiles: hh <= 128
constexpr int LROWS = 64;              // packed rows per ladder block (16 per wave)
constexpr int WROWS = 16;              // packed rows per wave tile (T X term)
constexpr int kGasTableMax = 4096;     // entries of the generated-V gas table (LDS, 32 KB)

// X = V^hhhh + sum_{c<d} V^hhpp[kl][cd] tas[cd][ij], split over k segments of
// KXSEG packed pairs: a block computes one 16x16 tile of one segment
// (tile.row0 = first row kl, tile.aux = first column ij, tile.nrows =
// segment), so that every wave has at most two k chunks in its dependent
// chain; the four waves take every fourth chunk and their partial tiles are
// added in wave order through LDS.  A channel of one segment writes X
// (V^hhhh added) directly; the others write their segments' partial blocks
// (xp), which k_fold_x (kernels_update.hpp) adds in segment order.  (A last-block fold inside
// this kernel needs device-scope fences, which write back the L2 of the
// XCD: X took 167 us instead of 14.)  FP64 arithmetic; tas in TT (FP32 in
// the FP32 iterations, whose ladder reads it so).
template <typename TT>
__global__ __launch_bounds__(kThreads) void k_ladder_x(const DTile *__restrict__ tiles,
                                                       const DChan *__restrict__ chans,
                                                       const double *__restrict__ vtp,
                                                       const TT *__restrict__ tas,
                                                       const double *__restrict__ vh,
                                                       const double *__restrict__ zero,
                                                       double *__restrict__ x,
                                                       double *__restrict__ xp,
                                                       const LoopCtl *__restrict__ ctl,
                                                       const LoopCtl *__restrict__ pctl) {
    if (ctl->done) return;
    __shared__ double part[kThreads / 64][64][4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    // the tiles of a channel (which read the same rows) on one XCD
    const DTile tl = tiles[xcd_slot(blockIdx.x, gridDim.x)];
    const DChan ch = chans[tl.chan];
    if (block_solve_stopped(pctl, ch.prob)) return;   // a stopped solve of a batch
    const int H = ch.hh, kl0 = tl.row0, ij0 = tl.aux, seg = tl.nrows;
    const int kend = min(ch.npk, (seg + 1) * KXSEG);
    const bool aok = kl0 + m < H, bok = ij0 + m < H;
    const TT *zt = reinterpret_cast<const TT *>(zero);
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = seg * KXSEG + wave * KC; k0 < kend; k0 += KC * (kThreads / 64)) {
        // every load unconditional (elements outside the block read the zero
        // buffer), so that the 16 loads of a chunk are in flight together
        double a[8], b[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int w = k0 + kq * 8 + s;
            const bool ok = w < kend;
            a[s] = *((ok && aok) ? vtp + ch.off_tp + (int64_t)w * H + kl0 + m : zero);
            b[s] = (double)*((ok && bok) ? tas + ch.off_tp + (int64_t)w * H + ij0 + m : zt);
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
                const int64_t o = (int64_t)row * H + col;
                if (ch.nxs == 1) x[ch.off_hhhh + o] = vh[ch.off_hhhh + o] + s;
                else xp[ch.off_xp + (int64_t)seg * H * H + o] = s;
            }
        }
    }
}

// Where the pppp ladder takes V^pppp from.
enum LadderSource {
    LV_STORED = 0,    // packed blocks in HBM (the stored-interaction mode)
    LV_GAS = 1,       // electron gas, generated per element from a table in LDS
    LV_GENERIC = 2,   // any generator, per element through gen_tbme
    LV_GAS_BATCH = 3, // a batch: the electron-gas form with each solve's table in
                      // LDS (staged when a block moves on to another solve), the
                      // generic generator for the channels of other solves
};
constexpr bool gas_source(int src) { return src == LV_GAS || src == LV_GAS_BATCH; }

// Per-element generation of packed V^pppp (the matrix-free mode, the
// calcFree* twins of src/calc.cpp, which rebuild V per channel with calcTBME).
// Entries are indexed like the row-pair lists (DChan::pr_off + packed index).
// Electron gas: inside a channel ka+kb = kc+kd and sa+sb = sc+sd, so
//   <ab||cd> = [sa==sc] T(ka-kc) - [sb==sc] T(kb-kc),
// with T the direct term, even in each component of the momentum transfer.
// A channel with pairs has either one particle of each spin in every pair
// (total spin 0: one term survives, <ab||cd> = s(ab) s(cd) T(k_up(ab) -
// k_up(cd)), s = +-1 by which particle is the up one) or two of one spin
// (both terms, no spin test).  The records hold momenta pre-scaled into
// 16-bit fields, (x | y << 16, z), so that two v_sad_u16 give
//   sum of |rec(a) - rec(c)| over the fields = byte offset of T(|ka-kc|)
// in the table (8-byte entries: rec1, rec2; 4-byte: rec1f, rec2f).
// Total spin 0: rec1 = k_up, rec2 = (sign mask of s, 0).  Same spin: rec1 and
// rec2 = the momenta of the first and the second state.
// Generic: rec1 = (first state, second state).
struct LadderGen {
    const int2 *rec1;
    const int2 *rec2;
    const int2 *rec1f;   // gas: rec1 scaled for the FP32 table (4-byte entries)
    const int2 *rec2f;   // gas, same-spin channels: rec2 scaled for the FP32 table
    const int2 *rec1w;   // gas, total-spin-0 channels: rec1f with the sign field of the table [T, -T]
    const double *tab;   // gas: T over (|dx|,|dy|,|dz|), row-major, then one 0
    int ntab;            // gas: entries including the final 0 (a batch: the largest solve's)
    Gen G;               // generic
    // a batch: per solve, (offset in tab, entries) of its gas table ((0, 0)
    // for a solve whose channels use the generic generator); nullptr for one
    // problem, whose table is tab[0 .. ntab)
    const int2 *ptab;
    const LoopCtl *pctl;   // a batch: per-solve loop states (stopped solves are skipped)
};

// The gas table into LDS in the arithmetic type: FP64 iterations T, FP32
// ones [T, -T] with every entry rounded (n entries each).
template <typename CT>
__device__ __forceinline__ void stage_gas_table(const double *tab, int n, void *lds) {
    CT *t = reinterpret_cast<CT *>(lds);
    for (int i = threadIdx.x & (kThreads - 1); i < n; i += kThreads) {
        const CT v = (CT)tab[i];
        t[i] = v;
        if (sizeof(CT) == 4) t[n + i] = -v;
    }
}

// Start of a ladder tile in a batch: false when the tile's solve has
// stopped; with per-solve gas tables, the table of the tile's solve goes to
// LDS when the block moves on to another solve (the block's previous tile
// ended with a barrier after its last table read).  prob / stopped cache
// the last solve seen.  Block-uniform; every thread calls it.
template <typename CT, int SRC>
__device__ __forceinline__ bool ladder_tile_begin(const LadderGen &lg, const DChan &ch, int &prob, bool &stopped,
                                                  void *tab_lds) {
    if constexpr (SRC == LV_GAS) return true;   // one problem, its table staged once
    if (!lg.pctl && SRC != LV_GAS_BATCH) return true;
    if (ch.prob != prob) {
        prob = ch.prob;
        stopped = block_solve_stopped(lg.pctl, prob);
        if (SRC == LV_GAS_BATCH && !stopped) {
            const int2 pt = lg.ptab[prob];
            if (pt.y > 0) {
                __syncthreads();
                stage_gas_table<CT>(lg.tab + pt.x, pt.y, tab_lds);
                __syncthreads();
            }
        }
    }
    return !stopped;
}

typedef __attribute__((address_space(3))) const double lds_double;

// LDS address of T(|k_up(ab) - k_up(cd)|) for the records of a total-spin-0
// channel (no spin field: the address stays inside the table).
__device__ __forceinline__ uint32_t gas_addr1(int2 r, int2 c, uint32_t base) {
    return __builtin_amdgcn_sad_u16((uint32_t)r.x, (uint32_t)c.x,
                                    __builtin_amdgcn_sad_u16((uint32_t)r.y, (uint32_t)c.y, base));
}

// v with its sign bit flipped where the mask m (0 or 0x80000000) says so.
__device__ __forceinline__ float flip_sign(float v, uint32_t m) { return __uint_as_float(__float_as_uint(v) ^ m); }
__device__ __forceinline__ double flip_sign(double v, uint32_t m) {
    return __hiloint2double(__double2hiint(v) ^ (int)m, __double2loint(v));
}

__device__ __forceinline__ double lds_load(uint32_t addr) {
    return *reinterpret_cast<lds_double *>(static_cast<uintptr_t>(addr));
}

typedef __attribute__((address_space(3))) const float lds_float;
__device__ __forceinline__ float lds_loadf(uint32_t addr) {
    return *reinterpret_cast<lds_float *>(static_cast<uintptr_t>(addr));
}

// Value of lane S of this lane's 16-lane row (DPP row_newbcast).
template <int S>
__device__ __forceinline__ int row_bcast(int v) {
    return __builtin_amdgcn_mov_dpp(v, 0x150 + S, 0xF, 0xF, false);
}

template <int S, int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (S < N) {
        f(std::integral_constant<int, S>{});
        static_for<S + 1, N>(f);
    }
}

// D_p[u][ij] = sum_w Vp[u][w] tas[w][ij] for 64 packed rows u of one channel
// per tile (16 per wave) and NT*16 >= hh columns, split over k segments of
// KSEG packed pairs (tile.aux = segment): segment s writes its partial sum to
// slab s of the channel's D_p; k_ladder_tx writes 1/2 tas X to slab nseg, and
// the update adds the slabs in order.  The k range of a tile is one
// sequence of KCH-wide chunks.  In a chunk, k-slot q of MFMA step s is
// k = 8 (s/2) + 2q + s%2: the four lanes of a row read 64 contiguous bytes per
// 16-byte load, so a load instruction touches half cache lines whose other
// halves the next instruction reads (rows are 128-byte aligned).  Software
// pipeline: while the MFMAs of chunk c run, the A values of chunk c+1 (HBM)
// and its B tile (L2-resident amplitudes / X) are loaded into registers; the
// B tile then goes to the other LDS buffer, one barrier per chunk.  With a
// generated V, lane j of 16-lane row q holds the record of k-slot (q, j) of
// the chunk (G <= 16), prefetched one chunk ahead; MFMA step s broadcasts
// lane s of each row, and each lane builds its A value from that record and
// its row's records.  Every source gives the same values in the same
// summation order, so the modes agree bit for bit.  Block b runs the tiles
// bptr[b] .. bptr[b+1]-1 (the generated modes launch fewer blocks than tiles,
// so the gas table is staged into LDS once per block).
//
// Arithmetic type CT: double, or float for the mixed-precision iterations
// (V^pppp stored in FP32, or generated in FP64 and rounded; the antisymmetrised
// amplitudes stored in FP32 by the update, X rounded on load; FP32
// accumulation; D_p written in FP32, which holds the FP32 sums exactly).  For
// float the
// k-slot order is k = 16 (s/4) + 4q + s%4, again 64 contiguous bytes per row
// and load instruction.
template <typename CT>
__device__ __forceinline__ int kslot(int q, int s) {
    return sizeof(CT) == 8 ? 8 * (s >> 1) + 2 * q + (s & 1) : 16 * (s >> 2) + 4 * q + (s & 3);
}

template <int NT>
struct LadderShape {
    // k per pipeline stage: the B tile (KCH x 16 NT, double-buffered) stays
    // within 40 KB of LDS
    static constexpr int KCH = NT <= 5 ? 32 : 16;
    // packed pairs per k segment: whole channels for the narrow classes (many
    // tiles, one prologue each), 256 for the few wide channels
    static constexpr int KSEG = NT <= 2 ? 1024 : 256;
};

// The kernels of the pppp ladder a class can run on (the plan's kernel kind).
enum LadderKind {
    LK_DEFAULT = 0,   // FP32 iterations: the one-row-per-lane kernel for class 1, the
                      // 16x16x4 one for the others; FP64: the 16x16x4 one
    LK_MFMA16 = 1,    // v_mfma_{f32,f64}_16x16x4 (ladder_run) in every class
    LK_BF16X3 = 2,    // FP32 iterations: BF16x3 split products, v_mfma_f32_16x16x32_bf16
    LK_BF16X3SYM = 3, // FP32 iterations, class 1: BF16x3 on the symmetry of V^pppp (wave_k_tile)
    LK_BF16X3W = 4,   // FP32 iterations, class 1: BF16x3, waves split k, four rows per lane (wave_k_tile)
};


// B-tile LDS of ladder_run: two buffers of the chunk's KCH x 16 NT
// amplitudes.  Split products hold them as BF16 high and low parts, column
// by column: column c of buffer b, part h (0 high, 1 low) holds the 32
// k-slots of the chunk in MFMA order (lane group q, element e at 8q + e) at
// BF16 offset ((2 b + h) W + c) CSTR, 80-byte columns (16-byte reads of
// neighbouring columns fall on different banks).
template <typename CT, int NT, bool SPLIT>
struct LadderTile {
    static constexpr int KCH = SPLIT ? 32 : LadderShape<NT>::KCH;
    static constexpr int W = NT * 16;
    // row stride (elements) of the unsplit tile: an MFMA step reads rows
    // kslot(kq, s) for the four lane groups kq; for FP32 the two groups of a
    // half-wave then sit 16 banks apart (4 STR = 16 mod 32), so they do not
    // conflict
    static constexpr int STR = sizeof(CT) == 8 ? W + 1 : W + 4;
    static constexpr int CSTR = 40;
    static constexpr size_t BYTES = SPLIT ? (size_t)2 * 2 * W * CSTR * 2 : (size_t)2 * KCH * STR * sizeof(CT);
};

// Bytes of the gas table at the start of the ladder's dynamic LDS: FP64
// iterations hold T in FP64, FP32 ones [T, -T] in FP32 (the one-row-per-lane
// kernel reads both halves, the others the first); 16-byte aligned.
__host__ __device__ constexpr size_t ladder_tab_bytes(int ntab) { return ((size_t)8 * ntab + 15) / 16 * 16; }

// Tiles b0 .. b1-1 of one class (NT) by one block; the gas table is already
// in LDS at byte address tbase, the B tiles go to lds_b
// (LadderTile<CT, NT, SPLIT>::BYTES).
//
// SPLIT (CT = float): BF16x3 split products.  Every FP32 operand x is split
// into x_h = bf16(x) and x_l = bf16(x - x_h) (round to nearest even), and
// a b is taken as a_l b_h + a_h b_l + a_h b_h, three
// v_mfma_f32_16x16x32_bf16 per 32-k chunk and 16 columns with FP32
// accumulation (the a_l b_l term, below 2^-16 |a b|, is dropped), against
// eight v_mfma_f32_16x16x4f32.  Lane (q, m) of the A operand holds k-slots
// (q, 0..7) of row m, the values the FP32 path uses in its eight MFMA steps;
// the B tile is split once as it goes to LDS.
template <typename CT, int NT, int SRC, bool SPLIT>
__device__ __forceinline__ void ladder_run(const DTile *__restrict__ tiles, int b0, int b1,
                                           const DChan *__restrict__ chans, const CT *__restrict__ vp,
                                           const LadderGen &lg, const CT *__restrict__ tas,
                                           const double *__restrict__ zero, CT *__restrict__ dp,
                                           unsigned char *lds_b, uint32_t tbase, void *tab_lds) {
    typedef LadderTile<CT, NT, SPLIT> LT;
    static_assert(!SPLIT || sizeof(CT) == 4, "split products from FP32 operands");
    constexpr int KCH = LT::KCH, KSEG = LadderShape<NT>::KSEG, G = KCH / 4;
    constexpr int W = LT::W, STR = LT::STR, CSTR = LT::CSTR;
    constexpr int BPT = (KCH * W + kThreads - 1) / kThreads;
    constexpr bool GEN = SRC != LV_STORED;
    constexpr int VW = 16 / sizeof(CT);   // elements per 16-byte load of stored V
    typedef typename Mma<CT>::acc Acc;
    typedef CT vec __attribute__((ext_vector_type(VW)));
    static_assert(G <= 16, "one k-slot record per lane of a 16-lane row");
    static_assert(KSEG % KCH == 0 && G % VW == 0, "k segments of whole chunks, chunks of whole 16-byte loads");
    static_assert(!SPLIT || (G == 8 && BPT == 2 * NT), "split: 32-k chunks, one MFMA k-group per lane group");
    CT(*Bs)[KCH][STR] = reinterpret_cast<CT(*)[KCH][STR]>(lds_b);
    __bf16 *Bsp = reinterpret_cast<__bf16 *>(lds_b);
    // tid < kThreads (the launch bound) as a value the compiler can see, so that
    // the B-tile slot bounds (KCH * W is a multiple of kThreads) fold away and
    // the tile loads stay ahead of their stores, outside any branch
    const int tid = threadIdx.x & (kThreads - 1);
    const int wave = tid >> 6, lane = tid & 63, m = lane & 15, kq = lane >> 4;
    // FP32 iterations read the FP32 table through the records scaled for it
    constexpr bool F32 = sizeof(CT) == 4;
    const int2 *grec1_ = F32 ? lg.rec1f : lg.rec1, *grec2 = F32 ? lg.rec2f : lg.rec2;

    int prob = -1;
    bool stopped = false;
    for (int ti = b0; ti < b1; ti++) {
        const DTile tl = tiles[ti];
        const DChan ch = chans[tl.chan];
        if (!ladder_tile_begin<CT, SRC>(lg, ch, prob, stopped, tab_lds)) continue;
        // MODE 1 (ONE): a total-spin-0 channel of the generated electron gas,
        // V from one table term (records of k_up; the signs s(cd) applied to
        // the B rows, s(ab) to the output rows, so every product equals the
        // two-term one).  MODE 2 (SAME): pairs of one spin, two terms without
        // the spin mask (same addresses, same values).
        auto tile = [&](auto mode_) {
            constexpr int MODE = decltype(mode_)::value;
            constexpr bool ONE = MODE == 1, SAME = MODE == 2;
            // Total-spin-0 signs: the unsplit path folds s(cd) into the B rows
            // and s(ab) into the output rows (exact: the FP32 MFMA rounds
            // symmetrically).  The BF16 MFMA's sums are not symmetric under a
            // sign flip of all products, so the split path builds the signed
            // element itself from the table [T, -T] through the records with
            // the sign field (rec1w), like the one-row-per-lane kernel: stored
            // and generated V then feed the same products.
            constexpr bool FOLD = ONE && !SPLIT;
            constexpr bool GAS = gas_source(SRC) && MODE != 0;   // else the generic generator
            const int2 *grec1 = SPLIT && ONE ? lg.rec1w : grec1_;
            const int H = ch.hh, npk = ch.npk, ldv = ch.ldv, seg = tl.aux;
            const int kbeg = seg * KSEG, kend = min(npk, kbeg + KSEG);
            const int nv = (kend - kbeg + KCH - 1) / KCH;
            const int lrow = wave * 16 + m;
            const bool rok = lrow < tl.nrows;
            const int64_t u = tl.row0 + lrow;
            const CT *av = GEN ? nullptr : vp + ch.off_vp + u * ldv;
            int2 ra = {0, 0}, rb = {0, 0};
            if (GEN) {
                const int2 *z2 = reinterpret_cast<const int2 *>(zero);
                ra = *(rok ? (GAS ? grec1 : lg.rec1) + ch.pr_off + u : z2);
                if (GAS && SAME) rb = *(rok ? grec2 + ch.pr_off + u : z2);
            }

            // Every load is unconditional: an element outside the block reads the
            // zero buffer (stored V) or lies outside its buffer and reads 0 (the
            // buffer loads below), so no load sits behind a branch and the
            // prefetches stay in flight across the MFMAs.
            auto load_a = [&](int v, CT (&a)[G]) {   // stored V: V chunk v
                const int kc0 = kbeg + v * KCH;
#pragma unroll
                for (int h = 0; h < G / VW; h++) {
                    // k-slots (kq, VW h .. VW h + VW-1) are VW consecutive k; k and
                    // ldv are multiples of VW, so the group lies inside the row
                    const int k = kc0 + kslot<CT>(kq, VW * h);
                    const vec w = *reinterpret_cast<const vec *>((rok && k < kend && k < ldv)
                                                                     ? av + k
                                                                     : reinterpret_cast<const CT *>(zero));
#pragma unroll
                    for (int e = 0; e < VW; e++) a[VW * h + e] = w[e];
                }
            };
            // Buffer loads of the records (generated V) and of the B tiles: rows
            // at or past kend lie outside their buffers and read 0.
            const BufRsrc rsrc =
                GEN ? make_rsrc((GAS ? grec1 : lg.rec1) + ch.pr_off, (uint32_t)kend * 8u) : make_rsrc(zero, 0);
            const uint32_t roff = m < G ? (uint32_t)(kbeg + kslot<CT>(kq, m)) * 8u : kOOB;
            auto load_r = [&](int v, int2 &rr) {  // generated V: record of k-slot (kq, m) of V chunk v
                rr = buf_i2(rsrc, roff + (uint32_t)(v * KCH) * 8u);
            };
            auto gen_a = [&](int2 rr, CT (&a)[G]) {  // generated V from the chunk's records
                static_for<0, G>([&](auto S) {
                    constexpr int s = decltype(S)::value;
                    const int2 r = {row_bcast<s>(rr.x), row_bcast<s>(rr.y)};
                    if (F32 && GAS && ONE) {
                        a[s] = lds_loadf(gas_addr1(ra, r, tbase));
                    } else if (F32 && GAS) {
                        a[s] = lds_loadf(gas_addr1(ra, r, tbase)) - lds_loadf(gas_addr1(rb, r, tbase));
                    } else if (GAS && ONE) {
                        a[s] = (CT)lds_load(gas_addr1(ra, r, tbase));
                    } else if (GAS) {
                        a[s] = (CT)(lds_load(gas_addr1(ra, r, tbase)) - lds_load(gas_addr1(rb, r, tbase)));
                    } else if (F32) {
                        a[s] = gen_tbme_f32(lg.G, ra.x, ra.y, r.x, r.y);
                    } else {
                        a[s] = (CT)gen_tbme(lg.G, ra.x, ra.y, r.x, r.y);
                    }
                });
            };
            // B tile: k rows k0.. of the channel's npk x H amplitude rows
            // (ONE: with the sign mask of row k, applied when the tile is
            // stored).  Unsplit: element q of a thread is slot idx = tid + 256 q
            // of the KCH x W tile.  Split: elements 2q and 2q+1 are rows 2p and
            // 2p + 1 of column c for the pair slot tid + 256 q = p W + c, so a
            // thread stores the adjacent k-slots of both rows as one 32-bit word
            // per part.
            const BufRsrc bsrc = make_rsrc(tas + ch.off_tp, (uint32_t)kend * (uint32_t)H * (uint32_t)sizeof(CT));
            const BufRsrc ssrc = FOLD ? make_rsrc(lg.rec2 + ch.pr_off, (uint32_t)kend * 8u) : make_rsrc(zero, 0);
            uint32_t boff[BPT], soff[BPT];
#pragma unroll
            for (int q = 0; q < BPT; q++) {
                int kk, col;
                bool ok;
                if (SPLIT) {
                    const int idx = tid + kThreads * (q >> 1);
                    kk = 2 * (idx / W) + (q & 1);
                    col = idx - (idx / W) * W;
                    ok = true;
                } else {
                    const int idx = tid + kThreads * q;
                    kk = idx / W;
                    col = idx - kk * W;
                    ok = idx < KCH * W;
                }
                boff[q] = (ok && col < H) ? (uint32_t)(kk * H + col) * (uint32_t)sizeof(CT) : kOOB;
                soff[q] = ok ? (uint32_t)kk * 8u : kOOB;
            }
            auto load_b = [&](int k0, CT (&b)[BPT], uint32_t (&sm)[BPT]) {
                const uint32_t ob = (uint32_t)k0 * (uint32_t)H * (uint32_t)sizeof(CT), os = (uint32_t)k0 * 8u;
#pragma unroll
                for (int q = 0; q < BPT; q++) {
                    b[q] = buf_load<CT>(bsrc, boff[q] + ob);
                    if (FOLD) sm[q] = buf_u32(ssrc, soff[q] + os);
                }
            };
            auto store_b = [&](int buf, const CT (&b)[BPT], const uint32_t (&sm)[BPT]) {
                if constexpr (SPLIT) {
#pragma unroll
                    for (int q = 0; q < BPT / 2; q++) {
                        const int idx = tid + kThreads * q, p = idx / W, col = idx - p * W;
                        // k-slot of row 2p: lane group ((2p) % 16) / 4, element
                        // 4 ((2p) / 16) + (2p) % 4 (row 2p + 1 is the next element)
                        const int k = 2 * p, pos = ((k & 15) >> 2) * 8 + 4 * (k >> 4) + (k & 3);
                        const u32x2 hl = split_pair(b[2 * q], b[2 * q + 1]);
                        const uint32_t hw = hl[0], lw = hl[1];
                        __bf16 *ph = Bsp + ((2 * buf) * W + col) * CSTR + pos;
                        __bf16 *pl = Bsp + ((2 * buf + 1) * W + col) * CSTR + pos;
                        *reinterpret_cast<uint32_t *>(ph) = hw;
                        *reinterpret_cast<uint32_t *>(pl) = lw;
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < BPT; q++) {
                        const int idx = tid + kThreads * q, kk = idx / W, col = idx - kk * W;
                        if (idx < KCH * W) Bs[buf][kk][col] = FOLD ? flip_sign(b[q], sm[q]) : b[q];
                    }
                }
            };

            Acc acc[NT];
#pragma unroll
            for (int n = 0; n < NT; n++) acc[n] = Acc{};
            auto mfma_chunk = [&](int buf, const CT (&a)[G]) {
                if constexpr (SPLIT) {
                    u32x4 hv, lv;
#pragma unroll
                    for (int j = 0; j < G / 2; j++) {
                        const u32x2 hl = split_pair(a[2 * j], a[2 * j + 1]);
                        hv[j] = hl[0];
                        lv[j] = hl[1];
                    }
                    const bf16x8 ah = __builtin_bit_cast(bf16x8, hv), al = __builtin_bit_cast(bf16x8, lv);
#pragma unroll
                    for (int n = 0; n < NT; n++) {
                        const __bf16 *col = Bsp + ((2 * buf) * W + n * 16 + m) * CSTR + 8 * kq;
                        const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(col);
                        const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(col + W * CSTR);
                        acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[n], 0, 0, 0);
                        acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[n], 0, 0, 0);
                        acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[n], 0, 0, 0);
                    }
                } else {
#pragma unroll
                    for (int s = 0; s < G; s++) {
#pragma unroll
                        for (int n = 0; n < NT; n++)
                            acc[n] = mfma16(a[s], Bs[buf][kslot<CT>(kq, s)][n * 16 + m], acc[n]);
                    }
                }
            };
            CT b0[BPT], b1[BPT];
            uint32_t sm0[BPT] = {}, sm1[BPT] = {};

            // V chunks, software-pipelined over two register sets and two LDS
            // buffers, unrolled by two so that no copy joins the sets: while the
            // MFMAs of chunk v run, chunk v+1 is in flight, and the wait for it
            // comes only before its own MFMAs, with chunk v+2 already issued.
            // The B tile of chunk v+2 is loaded into registers during chunk v and
            // goes to LDS at the end of chunk v+1, so each tile load has two chunks
            // of work behind it.  Loads past the last chunk lie outside their
            // buffers and read 0.  With a generated V the A values of chunk v+1 are
            // built between the (dependent) MFMAs of chunk v, from records loaded
            // two chunks ahead.
            CT a0[G], a1[G];
            int2 r0 = {0, 0}, r1 = {0, 0};
            auto stage = [&](int v, int buf, CT (&acur)[G], CT (&anext)[G], int2 &rnext, int2 &rnext2,
                             CT (&bload)[BPT], uint32_t (&smload)[BPT], const CT (&bnext)[BPT],
                             const uint32_t (&smnext)[BPT]) {
                load_b(kbeg + (v + 2) * KCH, bload, smload);
                if (!GEN) {
                    load_a(v + 1, anext);
                } else {
                    load_r(v + 2, rnext2);
                    gen_a(rnext, anext);
                }
                mfma_chunk(buf, acur);
                store_b(buf ^ 1, bnext, smnext);
                __syncthreads();
            };
            load_b(kbeg, b0, sm0);
            load_b(kbeg + KCH, b1, sm1);
            if (!GEN) {
                load_a(0, a0);
            } else {
                load_r(0, r0);
                load_r(1, r1);
                gen_a(r0, a0);
            }
            store_b(0, b0, sm0);
            __syncthreads();
            for (int v = 0; v < nv; v += 2) {
                stage(v, 0, a0, a1, r1, r0, b0, sm0, b1, sm1);
                if (v + 1 >= nv) break;
                stage(v + 1, 1, a1, a0, r0, r1, b1, sm1, b0, sm0);
            }
#pragma unroll
            for (int n = 0; n < NT; n++) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int row = wave * 16 + Mma<CT>::row(q, lane), col = n * 16 + m;
                    if (row < tl.nrows && col < H) {
                        CT val = acc[n][q];
                        if (FOLD)
                            val = flip_sign(val, *reinterpret_cast<const uint32_t *>(lg.rec2 + ch.pr_off + tl.row0 + row));
                        dp[ch.off_dp + (int64_t)seg * npk * H + (int64_t)(tl.row0 + row) * H + col] = val;
                    }
                }
            }
        };
        // a gas channel with pairs is total-spin-0 or same-spin (the host sets
        // spin0 so); in a batch spin0 = 0: the generic generator (other solves)
        if (gas_source(SRC) && ch.spin0 == 1) tile(std::integral_constant<int, 1>{});
        else if (SRC == LV_GAS || (SRC == LV_GAS_BATCH && ch.spin0 == 2)) tile(std::integral_constant<int, 2>{});
        else tile(std::integral_constant<int, 0>{});
    }
}

// ---- FP32 pppp ladder, channels with hh <= 16 NTW: one packed row per lane
//
// v_mfma_f32_4x4x1f32 runs 16 blocks of 4x4 outer products with k = 1: lane
// l, register r of the result holds A[lane 4(l/4) + r] * B[lane l] (+ C).
// With A_l = V[u0 + l][k] (one packed row per lane) and
// B_l = tas[k][4g + l%4] every block sees the same k, so a wave's 64 rows
// step through k together:
//   D[u0 + 4(l/4) + r][4g + l%4] += V[u0 + 4(l/4) + r][k] tas[k][4g + l%4],
// one MFMA per k and column group g of 4 columns, ceil(hh/4) groups (the
// channel's hh sets the MFMA work).  The k-slot records of a generated V are
// the same for the whole wave: scalar loads bring a chunk's records into
// SGPRs a chunk ahead, and a lane builds the element of its own row from
// them: two v_sad_u16 and one FP32 table read for a total-spin-0 channel,
// four, two reads and an FP32 subtraction for a same-spin one (the element of
// gen_tbme_f32), gen_tbme_f32 for other generators, a load for a stored V.
// Total-spin-0 channels read a table [T, -T]: the sign s(ab) s(cd) of the
// element is the distance of the records' sign fields (0 or the byte size
// of T), so the element comes signed from one read.
//
// A block runs the same tiles as ladder_run (64 rows x one k segment); its
// four waves take the tile's k chunks in turn (wave w: chunks w, w + 4, ...).
// Each wave stages its chunks' B rows (transposed, column-major: a lane reads
// 4 consecutive k of its column in one 16-byte read) in a double-buffered
// region of its own, loaded two chunks ahead; the V elements of its next
// chunk are built while the MFMAs of this one run.  Within a wave LDS
// accesses complete in order, so no barrier is needed inside the k loop.  At
// the end of a tile the four partial tiles are added in wave order through
// LDS.  The stored and generated modes give the same sums bit for bit (a
// sign flip is exact and commutes with every rounding).
constexpr int WKC = 16;        // k per chunk
constexpr int kRecPad = WKC;   // zero record entries past the last pair

template <int NTW>
struct WaveLadder {
    static constexpr int W = 16 * NTW;         // padded columns of the B tile
    static constexpr int NG = 4 * NTW;         // column groups of 4
    static constexpr int BSTR = WKC + 4;       // floats per transposed column (bank spread of the 16-byte reads)
    static constexpr int BUF = W * BSTR;       // floats per buffer
    static constexpr int PER = WKC * W / 64;   // B elements a lane stages per chunk
    // accumulator sets (k parity): two when few column groups would chain
    // dependent MFMAs back to back
    static constexpr int NACC = NTW == 1 ? 2 : 1;
    static constexpr int WREG = 2 * BUF;                // floats of a wave's region
    static constexpr int RED = 3 * NG * 4 * 64;         // floats of the partial tiles of waves 1-3
    // floats of the block's region after the gas table: the four wave
    // regions, which the partial tiles reuse at the end of a tile
    static constexpr int REGION = 4 * WREG > RED ? 4 * WREG : RED;
};

typedef __attribute__((address_space(4))) const int const_int;

template <int NTW, int SRC, int MODE>
__device__ __forceinline__ void wave_tile(const DTile &tl, const DChan &ch, const float *__restrict__ vp,
                                          const LadderGen &lg, const float *__restrict__ tas,
                                          float *__restrict__ dp, float *__restrict__ region, uint32_t tb32,
                                          int wave, int lane) {
    typedef WaveLadder<NTW> S;
    constexpr bool ONE = MODE == 1, SAME = MODE == 2, GEN = SRC != LV_STORED;
    constexpr bool GAS = gas_source(SRC) && MODE != 0;   // else the generic generator
    const int H = ch.hh, npk = ch.npk, seg = tl.aux;
    const int kbeg = seg * LadderShape<NTW>::KSEG, kend = min(npk, kbeg + LadderShape<NTW>::KSEG);
    const int ng = (H + 3) >> 2;
    const int nch = (kend - kbeg + WKC - 1) / WKC;
    float *bw = region + wave * S::WREG;
    // rows past the tile repeat its first row: finite values, never stored
    const int64_t u = tl.row0 + (lane < tl.nrows ? lane : 0);
    // k-slot records (uniform: scalar loads) and this lane's row records
    const int2 *rk = GAS ? (ONE ? lg.rec1w : lg.rec1f) : lg.rec1;
    int2 ra = {0, 0}, rb = {0, 0};
    if (GEN) ra = rk[ch.pr_off + u];
    if (GAS && SAME) rb = lg.rec2f[ch.pr_off + u];
    // the channel's records as constant memory (scalar loads for uniform addresses)
    const_int *er = (const_int *)(uintptr_t)(rk + ch.pr_off);
    const float *av = vp + ch.off_vp + u * ch.ldv;   // stored V: this lane's row (pad entries are 0)

    // staging of chunk c: B element idx = lane + 64 q is (k row idx / W,
    // column idx % W), rows at or past kend and columns past hh read 0
    const BufRsrc bsrc = make_rsrc(tas + ch.off_tp, (uint32_t)kend * (uint32_t)H * 4u);
    // (W is a power of two: kk = idx / W and col = idx % W are shifts and masks)
    static_assert((S::W & (S::W - 1)) == 0, "W a power of two");
    const int kk0 = lane / S::W, col = lane & (S::W - 1);   // element q: row kk0 + q 64 / W, column col
    const uint32_t boff0 = col < H ? (uint32_t)(kk0 * H + col) * 4u : kOOB;
    float breg[S::PER];
    auto load = [&](int c) {   // chunk c (its k rows may lie past kend: 0)
        const int k0 = kbeg + c * WKC;
#pragma unroll
        for (int q = 0; q < S::PER; q++)
            breg[q] = buf_load<float>(bsrc, boff0 + (uint32_t)(k0 + q * (64 / S::W)) * (uint32_t)H * 4u);
    };
    auto store = [&](int buf) {
        float *b = bw + buf * S::BUF + col * S::BSTR + kk0;
#pragma unroll
        for (int q = 0; q < S::PER; q++) b[q * (64 / S::W)] = breg[q];
    };
    // the WKC records of chunk c; past the last chunk those of row kend - 1
    // on (the kRecPad zero entries keep them inside the list), never used
    auto recs = [&](int c, int2 (&e)[WKC]) {
        if (GEN) {
            const_int *p = er + 2 * min(kbeg + c * WKC, kend - 1);
#pragma unroll
            for (int j = 0; j < WKC; j++) e[j] = int2{p[2 * j], p[2 * j + 1]};
        }
    };
    flt4 acc[S::NACC][S::NG];
#pragma unroll
    for (int g = 0; g < S::NG; g++)
#pragma unroll
        for (int h = 0; h < S::NACC; h++) acc[h][g] = flt4{0.f, 0.f, 0.f, 0.f};

    // V element of a row from a k-slot record e; same-spin channels return
    // the two terms (x1 - x2 is the element)
    auto term = [&](int2 e, float &a, float &x1, float &x2) {
        if (GAS && ONE) a = lds_loadf(gas_addr1(ra, e, tb32));
        else if (GAS) {
            x1 = lds_loadf(gas_addr1(ra, e, tb32));
            x2 = lds_loadf(gas_addr1(rb, e, tb32));
        } else a = gen_tbme_f32(lg.G, ra.x, ra.y, e.x, e.y);
    };
    auto mfma_quad = [&](int buf, int q, const float (&a)[WKC]) {
        const float *bb = bw + buf * S::BUF + (lane & 3) * S::BSTR + 4 * q;   // column 4 g + l % 4
#pragma unroll
        for (int g = 0; g < S::NG; g++) {
            if (g < ng) {   // uniform
                const flt4 b = *reinterpret_cast<const flt4 *>(bb + g * 4 * S::BSTR);
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[j % S::NACC][g] =
                        __builtin_amdgcn_mfma_f32_4x4x1f32(a[4 * q + j], b[j], acc[j % S::NACC][g], 0, 0, 0);
            }
        }
    };
    // the whole V chunk c (prologue)
    auto gen = [&](int c, const int2 (&e)[WKC], float (&a)[WKC]) {
        const int k0 = kbeg + c * WKC;
#pragma unroll
        for (int q = 0; q < WKC / 4; q++) {
            float x1[4], x2[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (SRC == LV_STORED) a[4 * q + j] = av[k0 + 4 * q + j];
                else term(e[4 * q + j], a[4 * q + j], x1[j], x2[j]);
            }
            if (GAS && SAME)
#pragma unroll
                for (int j = 0; j < 4; j++) a[4 * q + j] = x1[j] - x2[j];
        }
    };

    if (wave < nch) {   // uniform: this wave has chunks wave, wave + 4, ...
        float a0[WKC], a1[WKC];
        int2 e0[WKC], e1[WKC];
        recs(wave, e0);
        load(wave);
        store(0);
        load(wave + 4);
        recs(wave + 4, e1);
        gen(wave, e0, a0);
        // step: A of chunk c in acur, B of chunk c in buffer buf, the records
        // of chunk c + 4 in enext; the next chunk (c + 4) goes to the other
        // buffer and its A to anext while this one's MFMAs run, and the
        // records of chunk c + 8 are loaded into eload.  The order is kept by
        // scheduling barriers: per quarter of this chunk its MFMAs and the
        // table reads of a quarter of the next chunk, so that the LDS latency
        // sits behind MFMAs instead of in front of them.
        auto step = [&](int c, int buf, float (&acur)[WKC], float (&anext)[WKC], const int2 (&enext)[WKC],
                        int2 (&eload)[WKC]) {
            store(buf ^ 1);
            load(c + 8);
            const bool more = c + 4 < nch;   // uniform
            const int k4 = kbeg + (c + 4) * WKC;
#pragma unroll
            for (int q = 0; q < WKC / 4; q++) {
                float x1[4], x2[4];
                if (more) {
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        if (SRC == LV_STORED) anext[4 * q + j] = av[k4 + 4 * q + j];
                        else term(enext[4 * q + j], anext[4 * q + j], x1[j], x2[j]);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
                mfma_quad(buf, q, acur);
                __builtin_amdgcn_sched_barrier(0);
                if (more && GAS && SAME)
#pragma unroll
                    for (int j = 0; j < 4; j++) anext[4 * q + j] = x1[j] - x2[j];
            }
            recs(c + 8, eload);
        };
        for (int c = wave; c < nch; c += 8) {
            step(c, 0, a0, a1, e1, e0);
            if (c + 4 >= nch) break;
            step(c + 4, 1, a1, a0, e0, e1);
        }
    }
    // the four partial tiles, added in wave order: waves 1-3 park theirs
    // (register r of group g at [w - 1][g][r][lane]), wave 0 adds and stores
    __syncthreads();   // every wave is out of its k loop: the regions are free
    float *red = region;
    if (wave > 0) {
#pragma unroll
        for (int g = 0; g < S::NG; g++)
            if (g < ng)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float v = S::NACC == 2 ? acc[0][g][r] + acc[S::NACC - 1][g][r] : acc[0][g][r];
                    red[(((wave - 1) * S::NG + g) * 4 + r) * 64 + lane] = v;
                }
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = 4 * (lane >> 2) + r;
            const bool ok = row < tl.nrows;
            const int64_t o = ch.off_dp + (int64_t)seg * npk * H + (int64_t)(tl.row0 + (ok ? row : 0)) * H;
#pragma unroll
            for (int g = 0; g < S::NG; g++) {
                const int col = 4 * g + (lane & 3);
                if (g < ng) {
                    float v = S::NACC == 2 ? acc[0][g][r] + acc[S::NACC - 1][g][r] : acc[0][g][r];
#pragma unroll
                    for (int w = 1; w < 4; w++) v += red[(((w - 1) * S::NG + g) * 4 + r) * 64 + lane];
                    if (ok && col < H) dp[o + col] = v;
                }
            }
        }
    }
    __syncthreads();   // the regions are reused by the next tile
}

// ---- FP32 pppp ladder of class 1 on the symmetry of V^pppp (LK_BF16X3SYM)
//
// V^pppp of a channel is symmetric, <ab||cd> = <cd||ab> (real states; the
// generators and the stored blocks give V[P][Q] and V[Q][P] from the same
// operations), so a tile -- 64 packed rows P from row0 = 64 J -- builds
// only the elements V[P][Q] with Q >= row0 and uses each twice:
//  * D[P] += V[P][Q] T[Q], the tile's own rows, for every Q >= row0;
//  * D[Q] += V[P][Q] T[P] over the tile's rows, for Q past the tile
//    (Q >= row0 + 64): the part of row Q that the rows of tile J carry.
//    It goes to the channel's symmetric partial sums: the rows of row tile
//    J' = Q / 64 hold J' slabs of 64 x hh, slab J at
//    sym + ((J' (J' - 1) / 2 + J) 64 + Q - 64 J') hh; the update adds them
//    after the channel's D_p slabs.
// The first sum over the tile's rows and the second over the rows of the
// earlier tiles add up to row Q's sum over every k.  Elements built: the
// sum over the tiles of 64 (npk - row0), 0.52-0.57 of npk^2 for the
// electron gases at N=54 and N=66.
//
// Waves take the tile's 16-wide k chunks in turn (wave w: chunks w, w + 4,
// ...), so each wave's first chunk lies in the tile's diagonal block and
// every later one past it.  Lane (q, m) builds the elements of rows
// row0 + 16 rb + m (rb = 0..3) at k-slots k0 + 4 q + e (e = 0..3), from
// the records of the k-slots broadcast from lanes e of its 16-lane row (one
// broadcast serves four rows).  Both products are BF16x3 split products
// (x = x_h + x_l with x_h = bf16(x), x_l = bf16(x - x_h); a b = a_l b_h +
// a_h b_l + a_h b_h, FP32 accumulation):
//  * own rows, per row block: three v_mfma_f32_16x16x16_bf16 with
//    A[m][4q+e] = V[16 rb + m][k0 + 4 q + e], B[4q+e][n] = T[k0 + 4 q + e][n];
//  * rows of the chunk: the two BF16 parts of each element go, as 16-bit
//    words, into the wave's LDS blocks [k-slot][row] (one per part) and come
//    back transposed (lane (q, m) reads rows 32 g + 8 q .. + 7 of k-slot m,
//    16 bytes: the A operand of v_mfma_f32_16x16x32_bf16, B = the tile's own
//    T rows, split once per tile); 6 MFMAs per chunk, complete for the
//    chunk's 16 rows.
// Within a wave LDS accesses complete in order, so the transposition needs
// no barrier.  The own-row partial tiles of the four waves are added in
// wave order at the end of the tile.  The stored and generated modes use
// the same elements in the same order, so they agree bit for bit.  A tile
// of a later k segment (aux > 0: channels of more than one segment) only
// clears its slab of D_p: the segment-0 tile covers the whole k range.
//
// SYM = false (LK_BF16X3W): the same kernel without the symmetry -- every
// tile takes the k range of its segment (all of it from k = 0, like the
// other kinds) and only the own-row product, written to its segment's slab
// -- for the layout alone: one record broadcast per four rows (the
// 16x16x32 kernel broadcasts per row) and no partial sums for the update.
constexpr int SKC = 16;     // k-slots per chunk
constexpr int SPH = 72;     // 16-bit words per k-slot row of a wave's transposition blocks: 64 rows
                            // + 8 (the four row groups' writes 16 banks apart; the 16-byte reads
                            // of a row group cover all 64 banks once)

struct SymLadder {
    static constexpr int WBLK = 2 * SKC * SPH;     // 16-bit words of a wave's two blocks (high, low parts)
    static constexpr int RED = 4 * 4 * 4 * 64;     // floats of the four waves' own-row partial tiles
    static constexpr size_t BYTES = (size_t)4 * WBLK * 2 > (size_t)RED * 4 ? (size_t)4 * WBLK * 2 : (size_t)RED * 4;
};


template <int SRC, int MODE, bool SYM>
__device__ __forceinline__ void wave_k_tile(const DTile &tl, const DChan &ch, const float *__restrict__ vp,
                                         const LadderGen &lg, const float *__restrict__ tas,
                                         const double *__restrict__ zero, float *__restrict__ dp,
                                         float *__restrict__ dsym, unsigned char *lds, uint32_t tbase, int wave,
                                         int lane) {
    constexpr bool ONE = MODE == 1, SAME = MODE == 2, GEN = SRC != LV_STORED;
    constexpr bool GAS = gas_source(SRC) && MODE != 0;   // else the generic generator
    const int H = ch.hh, npk = ch.npk, m = lane & 15, q = lane >> 4;
    if (SYM && tl.aux > 0) {   // block-uniform
        float *o = dp + ch.off_dp + (int64_t)tl.aux * npk * H + (int64_t)tl.row0 * H;
        for (int l = wave * 64 + lane; l < tl.nrows * H; l += kThreads) o[l] = 0.f;
        return;
    }
    constexpr int KSEG = LadderShape<1>::KSEG;
    const int row0 = tl.row0, J = row0 >> 6, seg = SYM ? 0 : tl.aux;
    const int kbeg = SYM ? row0 : seg * KSEG, kend = SYM ? npk : min(npk, kbeg + KSEG);
    const int nch = (kend - kbeg + SKC - 1) / SKC;
    // the records of the split path: total-spin-0 channels with the sign field
    const int2 *grec1 = GAS ? (ONE ? lg.rec1w : lg.rec1f) : lg.rec1;
    // rows of this lane; rows past the tile repeat its first row (finite
    // elements: their T rows read 0 in the second product, their own sums are
    // never stored)
    int2 ra[4], rb2[4];
    const float *av[4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        const int r = 16 * rb + m;
        const int64_t u = row0 + (r < tl.nrows ? r : 0);
        ra[rb] = GEN ? grec1[ch.pr_off + u] : int2{0, 0};
        rb2[rb] = GAS && SAME ? lg.rec2f[ch.pr_off + u] : int2{0, 0};
        av[rb] = vp + ch.off_vp + u * ch.ldv;
    }
    // buffer loads: records and T rows past kend (or columns past hh) read 0
    const BufRsrc rsrc = GEN ? make_rsrc(grec1 + ch.pr_off, (uint32_t)kend * 8u) : make_rsrc(zero, 0);
    const BufRsrc bsrc = make_rsrc(tas + ch.off_tp, (uint32_t)kend * (uint32_t)H * 4u);
    const uint32_t roff = m < 4 ? (uint32_t)(kbeg + 4 * q + m) * 8u : kOOB;   // record of k-slot 4 q + m
    const uint32_t boff = m < H ? (uint32_t)((kbeg + 4 * q) * H + m) * 4u : kOOB;   // T[kbeg + 4 q][m]
    const uint32_t cstep = (uint32_t)SKC * (uint32_t)H * 4u;
    // B of the second product: the tile's T rows 32 g + 8 q + e in column m
    bf16x8 tbh[2], tbl[2];
    const uint32_t toff = m < H ? (uint32_t)((row0 + 8 * q) * H + m) * 4u : kOOB;
#pragma unroll
    for (int g = 0; g < 2; g++) {
        float x[8];
#pragma unroll
        for (int e = 0; e < 8; e++) x[e] = SYM ? buf_load<float>(bsrc, toff + (uint32_t)((32 * g + e) * H) * 4u) : 0.f;
        u32x4 hv, lv;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const u32x2 hl = split_pair(x[2 * j], x[2 * j + 1]);
            hv[j] = hl[0];
            lv[j] = hl[1];
        }
        tbh[g] = __builtin_bit_cast(bf16x8, hv);
        tbl[g] = __builtin_bit_cast(bf16x8, lv);
    }
    flt4 acc[4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) acc[rb] = flt4{0.f, 0.f, 0.f, 0.f};
    uint16_t *blh = reinterpret_cast<uint16_t *>(lds) + wave * SymLadder::WBLK, *bll = blh + SKC * SPH;
    // the channel's symmetric partial sums (the host gives every class-1
    // channel of two or more row tiles a place; a chunk past the tile exists
    // only in those)
    const bool has_sym = SYM && ch.sym >= 0;
    float *sym = has_sym ? dsym + (int64_t)ch.sym * 1024 : dsym;

    if (wave < nch) {   // uniform
        int2 rc = GEN ? buf_i2(rsrc, roff + (uint32_t)(wave * SKC) * 8u) : int2{0, 0};
        float bc[4];
#pragma unroll
        for (int e = 0; e < 4; e++) bc[e] = buf_load<float>(bsrc, boff + (uint32_t)wave * cstep + (uint32_t)(e * H) * 4u);
        for (int c = wave; c < nch; c += 4) {
            // the next chunk's records and T rows (past the last chunk: 0)
            const int cn = c + 4;
            const int2 rn = GEN ? buf_i2(rsrc, roff + (uint32_t)(cn * SKC) * 8u) : int2{0, 0};
            float bn[4];
#pragma unroll
            for (int e = 0; e < 4; e++) bn[e] = buf_load<float>(bsrc, boff + (uint32_t)cn * cstep + (uint32_t)(e * H) * 4u);
            const int k0 = kbeg + c * SKC;
            float a[4][4];
            if constexpr (!GEN) {
                const bool kok = k0 + 4 * q < ch.ldv;
#pragma unroll
                for (int rb = 0; rb < 4; rb++) {
                    const flt4 w = *reinterpret_cast<const flt4 *>(kok ? av[rb] + k0 + 4 * q
                                                                       : reinterpret_cast<const float *>(zero));
#pragma unroll
                    for (int e = 0; e < 4; e++) a[rb][e] = w[e];
                }
            } else {
                static_for<0, 4>([&](auto E) {
                    constexpr int e = decltype(E)::value;
                    const int2 r = {row_bcast<e>(rc.x), row_bcast<e>(rc.y)};
#pragma unroll
                    for (int rb = 0; rb < 4; rb++) {
                        if (GAS && ONE) a[rb][e] = lds_loadf(gas_addr1(ra[rb], r, tbase));
                        else if (GAS) a[rb][e] = lds_loadf(gas_addr1(ra[rb], r, tbase)) - lds_loadf(gas_addr1(rb2[rb], r, tbase));
                        else a[rb][e] = gen_tbme_f32(lg.G, ra[rb].x, ra[rb].y, r.x, r.y);
                    }
                });
            }
            // BF16 pairs: element 2 j + i of a row in half i of word j
            u32x2 ah[4], al[4];
            const u32x2 b01 = split_pair(bc[0], bc[1]), b23 = split_pair(bc[2], bc[3]);
            const u32x2 bh = {b01[0], b23[0]}, bl = {b01[1], b23[1]};
#pragma unroll
            for (int rb = 0; rb < 4; rb++) {
                const u32x2 p01 = split_pair(a[rb][0], a[rb][1]), p23 = split_pair(a[rb][2], a[rb][3]);
                ah[rb] = u32x2{p01[0], p23[0]};
                al[rb] = u32x2{p01[1], p23[1]};
            }
#pragma unroll
            for (int rb = 0; rb < 4; rb++) {
                acc[rb] = mfma_bf16_k16(al[rb], bh, acc[rb]);
                acc[rb] = mfma_bf16_k16(ah[rb], bl, acc[rb]);
                acc[rb] = mfma_bf16_k16(ah[rb], bh, acc[rb]);
            }
            if (has_sym && c >= 4) {   // uniform: the chunk lies past the tile
#pragma unroll
                for (int rb = 0; rb < 4; rb++)
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int o = (4 * q + e) * SPH + 16 * rb + m;
                        const int sh = 16 * (e & 1);
                        blh[o] = (uint16_t)(ah[rb][e >> 1] >> sh);
                        bll[o] = (uint16_t)(al[rb][e >> 1] >> sh);
                    }
                flt4 tacc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int g = 0; g < 2; g++) {
                    const bf16x8 xh = *reinterpret_cast<const bf16x8 *>(blh + m * SPH + 32 * g + 8 * q);
                    const bf16x8 xl = *reinterpret_cast<const bf16x8 *>(bll + m * SPH + 32 * g + 8 * q);
                    tacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, tbh[g], tacc, 0, 0, 0);
                    tacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, tbl[g], tacc, 0, 0, 0);
                    tacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, tbh[g], tacc, 0, 0, 0);
                }
                // rows k0 + 4 q + r of row tile Jt, slab J
                const int Jt = J + (c >> 2);
                float *o = sym + ((int64_t)(Jt * (Jt - 1) / 2 + J) * 64 + 16 * (c & 3)) * H;
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (m < H && k0 + 4 * q + r < kend) o[(4 * q + r) * H + m] = tacc[r];
            }
            rc = rn;
#pragma unroll
            for (int e = 0; e < 4; e++) bc[e] = bn[e];
        }
    }
    // the own-row partial tiles, added in wave order: every wave parks its
    // four (row block rb, register r at [wave][rb][r][lane]), wave w adds row
    // block w and stores it
    __syncthreads();   // every wave is out of its k loop: the LDS is free
    float *red = reinterpret_cast<float *>(lds);
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int r = 0; r < 4; r++) red[((wave * 4 + rb) * 4 + r) * 64 + lane] = acc[rb][r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
        float v = red[((0 * 4 + wave) * 4 + r) * 64 + lane];
#pragma unroll
        for (int w = 1; w < 4; w++) v += red[((w * 4 + wave) * 4 + r) * 64 + lane];
        const int row = 16 * wave + 4 * q + r;
        if (row < tl.nrows && m < H) dp[ch.off_dp + (int64_t)seg * npk * H + (int64_t)(row0 + row) * H + m] = v;
    }
    __syncthreads();   // the LDS is reused by the next tile
}

// ---- the grouped pppp-ladder launch
//
// One persistent launch covers the classes LO..HI (channels with hh in
// (16 (n-1), 16 n] for class n): the plan gives each class a run of the
// launch's blocks, its kernel kind and its tile runs.  Block b belongs to
// class n for L.blk0[n] <= b < L.blk0[n+1] and runs the class's tiles
// bptr[L.ptr0[n] + r] .. bptr[L.ptr0[n] + r + 1] - 1 (r = b - L.blk0[n],
// tiles counted from L.tile0[n]).  The gas table is staged into LDS once per
// block; the B tiles of the block's class follow it in the same dynamic LDS,
// sized on the host for the largest class and kind of the launch.
struct LadderLaunch {
    int32_t tile0[NTMAX + 1];
    int32_t ptr0[NTMAX + 1];
    int32_t blk0[NTMAX + 2];
    int32_t kind[NTMAX + 1];
    float *dsym;   // symmetric partial sums of class 1 (LK_BF16X3SYM)
};

// Bytes of the B-tile LDS of class NT under kernel kind `kind` (CT: the
// iterations' arithmetic type).
template <typename CT, int NT>
__host__ __device__ constexpr size_t ladder_class_bytes(int kind) {
    if (sizeof(CT) == 4 && kind == LK_BF16X3) return LadderTile<float, NT, true>::BYTES;
    if (sizeof(CT) == 4 && NT == 1 && (kind == LK_BF16X3SYM || kind == LK_BF16X3W)) return SymLadder::BYTES;
    if (sizeof(CT) == 4 && NT == 1 && kind == LK_DEFAULT) return (size_t)WaveLadder<1>::REGION * sizeof(float);
    return LadderTile<CT, NT, false>::BYTES;
}

template <typename CT, int NT, int SRC>
__device__ __forceinline__ void ladder_class(int kind, const DTile *__restrict__ tiles, int b0, int b1,
                                             const DChan *__restrict__ chans, const CT *__restrict__ vp,
                                             const LadderGen &lg, const CT *__restrict__ tas,
                                             const double *__restrict__ zero, CT *__restrict__ dp,
                                             float *__restrict__ dsym, unsigned char *lds_b, uint32_t tbase,
                                             void *tab_lds) {
    if constexpr (std::is_same<CT, float>::value) {
        if (kind == LK_BF16X3) {
            ladder_run<float, NT, SRC, true>(tiles, b0, b1, chans, vp, lg, tas, zero, dp, lds_b, tbase, tab_lds);
            return;
        }
        if constexpr (NT == 1) {
            if (kind == LK_BF16X3SYM || kind == LK_BF16X3W) {
                const int tid = threadIdx.x & (kThreads - 1);
                const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
                int prob = -1;
                bool stopped = false;
                for (int ti = b0; ti < b1; ti++) {
                    const DTile tl = tiles[ti];
                    const DChan ch = chans[tl.chan];
                    if (!ladder_tile_begin<float, SRC>(lg, ch, prob, stopped, tab_lds)) continue;
                    auto run = [&](auto sym_) {
                        constexpr bool S = decltype(sym_)::value;
                        if (gas_source(SRC) && ch.spin0 == 1)
                            wave_k_tile<SRC, 1, S>(tl, ch, vp, lg, tas, zero, dp, dsym, lds_b, tbase, wave, lane);
                        else if (SRC == LV_GAS || (SRC == LV_GAS_BATCH && ch.spin0 == 2))
                            wave_k_tile<SRC, 2, S>(tl, ch, vp, lg, tas, zero, dp, dsym, lds_b, tbase, wave, lane);
                        else
                            wave_k_tile<SRC, 0, S>(tl, ch, vp, lg, tas, zero, dp, dsym, lds_b, tbase, wave, lane);
                    };
                    if (kind == LK_BF16X3SYM) run(std::true_type{});
                    else run(std::false_type{});
                }
                return;
            }
            if (kind == LK_DEFAULT) {
                const int tid = threadIdx.x & (kThreads - 1);
                const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
                int prob = -1;
                bool stopped = false;
                for (int ti = b0; ti < b1; ti++) {
                    const DTile tl = tiles[ti];
                    const DChan ch = chans[tl.chan];
                    if (!ladder_tile_begin<float, SRC>(lg, ch, prob, stopped, tab_lds)) continue;
                    float *region = reinterpret_cast<float *>(lds_b);
                    if (gas_source(SRC) && ch.spin0 == 1)
                        wave_tile<1, SRC, 1>(tl, ch, vp, lg, tas, dp, region, tbase, wave, lane);
                    else if (SRC == LV_GAS || (SRC == LV_GAS_BATCH && ch.spin0 == 2))
                        wave_tile<1, SRC, 2>(tl, ch, vp, lg, tas, dp, region, tbase, wave, lane);
                    else
                        wave_tile<1, SRC, 0>(tl, ch, vp, lg, tas, dp, region, tbase, wave, lane);
                }
                return;
            }
        }
    }
    ladder_run<CT, NT, SRC, false>(tiles, b0, b1, chans, vp, lg, tas, zero, dp, lds_b, tbase, tab_lds);
}

template <typename CT, int SRC, int LO, int HI>
__global__ __launch_bounds__(kThreads) void k_ladder_grp(LadderLaunch L, const DTile *__restrict__ tiles,
                                                         const int32_t *__restrict__ bptr,
                                                         const DChan *__restrict__ chans,
                                                         const CT *__restrict__ vp, LadderGen lg,
                                                         const CT *__restrict__ tas,
                                                         const double *__restrict__ zero, CT *__restrict__ dp,
                                                         const LoopCtl *__restrict__ ctl) {
    if (ctl->done) return;
    extern __shared__ double lg_lds[];
    const int ntab = gas_source(SRC) ? lg.ntab : 0;
    if (SRC == LV_GAS) {   // one problem: its table, once (a batch stages each solve's as it goes)
        stage_gas_table<CT>(lg.tab, ntab, lg_lds);
        __syncthreads();
    }
    const uint32_t tbase = (uint32_t)(uintptr_t)(lds_double *)lg_lds;   // LDS byte address
    unsigned char *lds_b = reinterpret_cast<unsigned char *>(lg_lds) + ladder_tab_bytes(ntab);
    const int b = blockIdx.x;
    int nt = LO;
    while (nt < HI && b >= L.blk0[nt + 1]) nt++;   // uniform
    const int r = b - L.blk0[nt];
    const DTile *tl = tiles + L.tile0[nt];
    const int b0 = bptr[L.ptr0[nt] + r], b1 = bptr[L.ptr0[nt] + r + 1];
    const int kind = L.kind[nt];
    static_for<LO, HI + 1>([&](auto N) {
        constexpr int n = decltype(N)::value;
        if (nt == n)
            ladder_class<CT, n, SRC>(kind, tl, b0, b1, chans, vp, lg, tas, zero, dp, L.dsym, lds_b, tbase, lg_lds);
    });
}

// 1/2 tas X, the hhhh quadratic term of D (calc.cpp:171-187 with X from
// k_ladder_x), for 16 packed rows of one channel per wave (tile.row0/nrows,
// all hh columns), written to slab nseg of the channel's D_p.  Waves are
// independent: operands come straight from global memory (tas rows and the
// small X block stay in L1/L2).
template <typename CT, int NT>
__device__ __forceinline__ void ladder_tx_wave(const DTile &tl, const DChan &ch, const CT *__restrict__ tas,
                                               const double *__restrict__ x, const double *__restrict__ zero,
                                               CT *__restrict__ dp) {
    typedef typename Mma<CT>::acc Acc;
    const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    const int H = ch.hh;
    const bool rok = m < tl.nrows;
    const CT *at = tas + ch.off_tp + (int64_t)(tl.row0 + m) * H;
    const CT *zc = reinterpret_cast<const CT *>(zero);
    const double *X = x + ch.off_hhhh;
    Acc acc[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) acc[n] = Acc{};
    // per 16-wide k step: every operand load first, then the MFMAs (loaded
    // inside the MFMA loop, each product waited for its own loads)
#pragma unroll
    for (int kt = 0; kt < NT; kt++) {
        CT av[4], bv[4][NT];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = kt * 16 + 4 * kq + j;
            const bool kok = k < H;
            av[j] = (CT)0.5 * *((rok && kok) ? at + k : zc);
#pragma unroll
            for (int n = 0; n < NT; n++) {
                const int col = n * 16 + m;
                bv[j][n] = (CT) * ((kok && col < H) ? X + k * H + col : zero);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
#pragma unroll
            for (int n = 0; n < NT; n++) acc[n] = mfma16(av[j], bv[j][n], acc[n]);
        }
    }
    CT *out = dp + ch.off_dp + (int64_t)ch.nseg * ch.npk * H;
#pragma unroll
    for (int n = 0; n < NT; n++) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int row = Mma<CT>::row(q, lane), col = n * 16 + m;
            if (row < tl.nrows && col < H) out[(int64_t)(tl.row0 + row) * H + col] = acc[n][q];
        }
    }
}

// WIDE instances take the tiles of channels with hh > 80 (NT = 6..8), the
// others NT = 1..5; the tile list holds the narrow channels first.
template <typename CT, bool WIDE>
__global__ __launch_bounds__(kThreads) void k_ladder_tx(const DTile *__restrict__ tiles, int ntiles,
                                                        const DChan *__restrict__ chans,
                                                        const CT *__restrict__ tas,
                                                        const double *__restrict__ x,
                                                        const double *__restrict__ zero, CT *__restrict__ dp,
                                                        const LoopCtl *__restrict__ ctl,
                                                        const LoopCtl *__restrict__ pctl) {
    if (ctl->done) return;
    const int it = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (it >= ntiles) return;
    const DTile tl = tiles[it];
    const DChan ch = chans[tl.chan];
    if (wave_solve_stopped(pctl, ch.prob)) return;   // a stopped solve of a batch
    const int nt = (ch.hh + 15) / 16;
    if (WIDE) {
        switch (nt) {
            case 6: ladder_tx_wave<CT, 6>(tl, ch, tas, x, zero, dp); break;
            case 7: ladder_tx_wave<CT, 7>(tl, ch, tas, x, zero, dp); break;
            default: ladder_tx_wave<CT, 8>(tl, ch, tas, x, zero, dp); break;
        }
    } else {
        switch (nt) {
            case 1: ladder_tx_wave<CT, 1>(tl, ch, tas, x, zero, dp); break;
            case 2: ladder_tx_wave<CT, 2>(tl, ch, tas, x, zero, dp); break;
            case 3: ladder_tx_wave<CT, 3>(tl, ch, tas, x, zero, dp); break;
            case 4: ladder_tx_wave<CT, 4>(tl, ch, tas, x, zero, dp); break;
            default: ladder_tx_wave<CT, 5>(tl, ch, tas, x, zero, dp); break;
        }
    }
}

// Number of k segments of a channel's pppp ladder.
inline int ladder_segments(int npk, int nt) {
    const int kseg = nt == 1   ? LadderShape<1>::KSEG
                     : nt == 2 ? LadderShape<2>::KSEG
                     : nt == 3 ? LadderShape<3>::KSEG
                     : nt == 4 ? LadderShape<4>::KSEG
                     : nt == 5 ? LadderShape<5>::KSEG
                     : nt == 6 ? LadderShape<6>::KSEG
                     : nt == 7 ? LadderShape<7>::KSEG
                               : LadderShape<8>::KSEG;
    return npk > 0 ? (npk + kseg - 1) / kseg : 1;
}
