// Device-side types and helpers shared by the CCD kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "ccd_amd.h"

constexpr int kThreads = 256;   // every kernel runs 4 wave64s per block
constexpr int kZeroDoubles = 32;  // size of the zero buffer masked-off loads read (256 B)
constexpr int kElogCap = 1024;    // energies logged by a device-run iteration loop
constexpr int kSolveBatch = 4;    // iterations queued between two reads of the loop state

// The workgroups of a launch are dispatched to the 8 XCDs in turn (block b to
// XCD (b + c) mod 8 for some c), and each XCD has an L2 of its own.  The work
// item of block b of an n-block grid such that each XCD takes a contiguous
// run of the n items: XCD x gets items x (n / 8) + min(x, n mod 8) onward, in
// block order, so that neighbouring items that read the same rows share an
// L2.  A bijection of [0, n).
__host__ __device__ __forceinline__ int xcd_slot(int b, int n) {
    const int x = b & 7, q = b >> 3, per = n >> 3, rem = n & 7;
    return x * per + (x < rem ? x : rem) + q;
}

// Raw (unswizzled) buffer loads.  A lane whose byte offset is at or beyond the
// buffer's size reads 0 without touching memory (the hardware range check), so
// an element outside a tile needs no select: its offset is kOOB (the sum of
// kOOB and an in-buffer offset below 2^31 stays out of range).  Buffers hold
// fewer than 2^31 bytes.
constexpr uint32_t kOOB = 0x80000000u;
typedef __amdgpu_buffer_rsrc_t BufRsrc;
// base and bytes are wave-uniform; the descriptor is built from lane 0's values
// so that it lives in scalar registers.
__device__ __forceinline__ BufRsrc make_rsrc(const void *base, uint32_t bytes) {
    const uint64_t p = reinterpret_cast<uintptr_t>(base);
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)p);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(p >> 32));
    return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void *>(hi << 32 | lo), (short)0,
                                             __builtin_amdgcn_readfirstlane((int)bytes), 0x00020000);
}
__device__ __forceinline__ double buf_f64(BufRsrc r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ int2 buf_i2(BufRsrc r, uint32_t off) {
    return __builtin_bit_cast(int2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ uint32_t buf_u32(BufRsrc r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}
template <typename T>
__device__ __forceinline__ T buf_load(BufRsrc r, uint32_t off) {
    if constexpr (sizeof(T) == 8) return buf_f64(r, off);
    else return __builtin_bit_cast(float, buf_u32(r, off));
}

struct DChan {          // active direct channel (pp > 0 and hh > 0)
    int64_t off_t;      // block offset in the pphh layout
    int64_t off_vp;     // packed V^pppp block: npk x ldv, rows/cols = pairs a<b
    int64_t off_tp;     // packed-row blocks (npk x hh): antisymmetrised T, D_p, V^hhpp rows
    int64_t off_hhhh;   // hh x hh block of V^hhhh / X
    int32_t pp, hh;
    int32_t pp_off, hh_off;  // first pair in the pp / hh pair lists
    int32_t npk, ldv;        // number of pairs a<b; leading dimension (multiple of 8)
    int32_t pr_off, ndiag;   // first entry in the row-pair lists; number of pairs (a,a)
    int64_t off_dp;     // ladder output: nseg + 1 slabs of npk x hh partial sums
    int32_t nseg;       // number of k segments of the pppp ladder (slab nseg: 1/2 T X)
    int32_t prob;       // solve the channel belongs to (0 outside a batch)
    int32_t spin0;      // generated electron-gas V: 1 when every pair has one particle of each
                        // spin (one table term per element), 2 when both have the same spin
                        // (two terms, no spin mask), see the pppp ladder
    int32_t hh2;        // columns of the hole pairs i != j: (i,j) at 2v, (j,i) at 2v + 1
                        // (the pairs (i,i) follow), so column IJ < hh2 has its partner at IJ ^ 1
    int64_t off_xp;     // the nxs k-segment partial blocks (hh x hh) of X = V^hhhh + V^hhpp T
    int32_t nxs;
    int32_t sym;        // class-1 channels of two or more 64-row tiles: offset of their symmetric
                        // partial sums of the pppp ladder (LK_BF16X3SYM), in units of 1024 floats;
                        // -1 for the others
};

struct DMod {           // active mod channel (ph > 0 and hp > 0)
    int64_t off_ph;     // block offset in the phhp layout (also V^A, hp x ph)
    int64_t off_b;      // offset of its hp x hp block of V^B
    int32_t ph, hp;
    int32_t ph_off, hp_off;
    int32_t prob;       // solve the channel belongs to (0 outside a batch)
    int32_t pad;
};

struct DTile {          // a run of rows (or row pairs) of one active direct channel
    int32_t chan, row0, nrows, aux;
};

struct DRow {           // one contiguous row of a layout, for the per-particle transposes
    int32_t goff;       // element offset in the layout
    int32_t len;
    int32_t loff;       // offset in the particle's LDS block
    int32_t pad;
};

// Device-side state of the iteration loop (src/main.cpp:491-519 run on the
// device).  k_reduce applies the reference's stop rule after each iteration;
// once `done` is set every kernel of later iterations returns at entry, so
// the host can queue iterations ahead and read the outcome in batches.
struct LoopCtl {
    int done;        // set once the loop would have stopped
    int count;       // iterations completed
    int reason;      // CCD_STOP_* once done
    int max_iters;   // halt once more than this many iterations have run
    double prev;     // energy of the previous iteration (100.0 at the start)
    double tol;      // continue while |dE| > tol
    int cap;         // entries of the energy log
    int idle;        // a solve of a batch: iterations of the batch since it stopped
};

// Whether solve `prob` of a batch has stopped (pctl: the per-solve loop
// states; nullptr outside a batch, or where nothing may be skipped).  The
// contraction kernels skip the work of a stopped solve: the update keeps its
// amplitudes and never reads their output.  The flag may be set while a
// kernel runs (the reductions of the previous iteration run on another
// stream), so one thread reads it and shares it through LDS: the answer is
// the same for the whole block.  Every thread of the block must call it.
__device__ __forceinline__ bool block_solve_stopped(const LoopCtl *pctl, int prob) {
    if (!pctl) return false;
    __shared__ int stop_flag;
    __syncthreads();   // the previous call's readers are done with stop_flag
    if (threadIdx.x == 0) stop_flag = __hip_atomic_load(&pctl[prob].done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return stop_flag != 0;
}

// The same for kernels whose waves are independent: one answer per wave.
__device__ __forceinline__ bool wave_solve_stopped(const LoopCtl *pctl, int prob) {
    if (!pctl) return false;
    return __builtin_amdgcn_readfirstlane(
               __hip_atomic_load(&pctl[prob].done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0;
}

// A batch of independent solves merged into one problem: every s.p. state
// carries its solve's index in qn[5 s + 4], and the solve's generator
// parameters are looked up per element.
constexpr int kGenBatch = 0x40;

struct GenPart {        // generator parameters of one solve of a batch
    int kind;
    int nmax;
    double g;
    int64_t tab_off;    // its direct-term table inside Gen::table
};

struct Gen {            // on-device interaction
    int kind;
    int nmax;
    double g;
    const int32_t *qn;
    const double *table;
    const GenPart *part;  // kind == kGenBatch: per-solve parameters
};

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef float flt2 __attribute__((ext_vector_type(2)));

// Fixed-order block reduction: wave butterfly, then the four wave sums in order.
// Every thread of the block must call it; the result is valid on thread 0.
__device__ __forceinline__ double block_sum(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
    return s;
}

// Minnesota <ij|v|kl> of pure neutron matter, not antisymmetrised
// (src/InfMatterSPBasis.cpp:361-449), in the host model's operation order
// (no contraction); pref = the three prefactors +-V_0X / L^3 (pi/kappa_X)^1.5.
__device__ __attribute__((noinline)) double minnesota_direct(const int32_t *qi, const int32_t *qj,
                                                             const int32_t *qk, const int32_t *ql, double Lb,
                                                             const double *pref) {
#pragma clang fp contract(off)
    if (qi[0] + qj[0] != qk[0] + ql[0] || qi[1] + qj[1] != qk[1] + ql[1] || qi[2] + qj[2] != qk[2] + ql[2])
        return 0.0;
    const double kappa_R = 1.487, kappa_T = 0.639, kappa_S = 0.465;
    double qSquared = 0.;
    for (int i = 0; i < 3; i++) {
        const double ki = 2 * M_PI * qi[i] / Lb, kj = 2 * M_PI * qj[i] / Lb;
        const double kk = 2 * M_PI * qk[i] / Lb, kl = 2 * M_PI * ql[i] / Lb;
        const double bra = 0.5 * (ki - kj), ket = 0.5 * (kk - kl), tr = bra - ket;
        qSquared += tr * tr;
    }
    const double V_R = pref[0] * exp(-qSquared / (4 * kappa_R));
    const double V_T = pref[1] * exp(-qSquared / (4 * kappa_T));
    const double V_S = pref[2] * exp(-qSquared / (4 * kappa_S));
    const double spinEx = (qi[3] == ql[3] && qj[3] == qk[3]) ? 1 : 0;
    const double isoSpinEx = (qi[4] == ql[4] && qj[4] == qk[4]) ? 1 : 0;
    const double ks = (qi[3] == qk[3] ? 1 : 0) * (qj[3] == ql[3] ? 1 : 0);
    const double kt = (qi[4] == qk[4] ? 1 : 0) * (qj[4] == ql[4] ? 1 : 0);
    const double IsIt = ks * kt, PsIt = spinEx * kt, PsPt = spinEx * isoSpinEx, IsPt = ks * isoSpinEx;
    return 0.5 * (V_R + 0.5 * V_T + 0.5 * V_S) * IsIt + 0.25 * (V_T - V_S) * PsIt -
           0.5 * (V_R + 0.5 * V_T + 0.5 * V_S) * PsPt - 0.25 * (V_T - V_S) * IsPt;
}

// Antisymmetrised <pq||rs> generated from quantum numbers; bitwise equal to
// the host model for the pairing model (src/PairingSPBasis.cpp:127-141) and
// the electron gas (src/ElectronGasSPBasis.cpp:419-471 with the direct term
// tabulated); infinite matter up to the device exp (src/InfMatterSPBasis.cpp:355-359).
// F32 = true: the element of the FP32 iterations -- for the electron gas its
// two direct terms rounded to FP32 and then subtracted, (float) d1 - (float) d2
// (one of them is 0 unless both pairs have one spin, so a total-spin-0 element
// is the FP64 one rounded); for the other generators the FP64 element rounded.
template <bool F32 = false>
__device__ inline auto gen_tbme_t(const Gen &G, int p, int q, int r, int s) {
    typedef typename std::conditional<F32, float, double>::type R;
    const int32_t *a = G.qn + 5 * p, *b = G.qn + 5 * q, *c = G.qn + 5 * r, *d = G.qn + 5 * s;
    int kind = G.kind, nmax = G.nmax;
    double g = G.g;
    const double *table = G.table;
    if (kind == kGenBatch) {   // the four states belong to one solve
        const GenPart P = G.part[a[4]];
        kind = P.kind; nmax = P.nmax; g = P.g; table = G.table + P.tab_off;
    }
    if (kind == CCD_GEN_INF_MATTER)
        return (R)(minnesota_direct(a, b, c, d, g, table) - minnesota_direct(a, b, d, c, g, table));
    if (kind == CCD_GEN_PAIRING) {
        double v = 0.0;
        if (a[0] == b[0] && c[0] == d[0] && a[1] == -b[1] && -c[1] == d[1]) {
            v = -0.5 * g;
            if (a[1] == -c[1]) v = -v;
        }
        return (R)v;
    }
    const bool cons = a[0] + b[0] == c[0] + d[0] && a[1] + b[1] == c[1] + d[1] && a[2] + b[2] == c[2] + d[2];
    if (!cons) return (R)0;
    const int m2 = 2 * nmax, w = 4 * nmax + 1;
    double d1 = 0.0, d2 = 0.0;
    if (a[3] == c[3] && b[3] == d[3])
        d1 = table[((a[0] - c[0] + m2) * w + (a[1] - c[1] + m2)) * w + (a[2] - c[2] + m2)];
    if (a[3] == d[3] && b[3] == c[3])
        d2 = table[((a[0] - d[0] + m2) * w + (a[1] - d[1] + m2)) * w + (a[2] - d[2] + m2)];
    return (R)d1 - (R)d2;
}
__device__ inline double gen_tbme(const Gen &G, int p, int q, int r, int s) { return gen_tbme_t<false>(G, p, q, r, s); }
__device__ inline float gen_tbme_f32(const Gen &G, int p, int q, int r, int s) { return gen_tbme_t<true>(G, p, q, r, s); }

// v_mfma_f64_16x16x4f64: A is 16x4 (lane l: row l%16, k-slot l/16), B is 4x16
// (lane l: k-slot l/16, column l%16), C is 16x16 with register q of lane l
// holding row 4q + l/16, column l%16.
__device__ __forceinline__ dbl4 mfma16(double a, double b, dbl4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x4f32: A and B as above; C register q of lane l holds row
// 4 (l/16) + q, column l%16.
__device__ __forceinline__ flt4 mfma16(float a, float b, flt4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// BF16x3 split products (x = x_h + x_l in bfloat16, a b ~ a_l b_h + a_h b_l +
// a_h b_h with FP32 accumulation).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef short sh4 __attribute__((ext_vector_type(4)));

// The BF16x3 parts of two FP32 values, each pair in one 32-bit word (the
// first value in the low half): [0] = bf16(x) (round to nearest even), [1] =
// bf16(x - [0]).  A pair per conversion (v_cvt_pk_bf16_f32 takes two
// values); the same numbers as converting one value at a time.
__device__ __forceinline__ u32x2 split_pair(float x0, float x1) {
    const flt2 x = {x0, x1};
    const uint32_t hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2));
    const flt2 h = {__uint_as_float(hi << 16), __uint_as_float(hi & 0xffff0000u)};
    return u32x2{hi, __builtin_bit_cast(uint32_t, __builtin_convertvector(x - h, bf16x2))};
}

// The parts of four values: [0] = the high parts, [1] = the low parts, as
// the operand of v_mfma_f32_16x16x16_bf16 (value e of the four in half e % 2
// of word e / 2).
struct Split4 {
    u32x2 h, l;
};
__device__ __forceinline__ Split4 split4(float x0, float x1, float x2, float x3) {
    const u32x2 p01 = split_pair(x0, x1), p23 = split_pair(x2, x3);
    return Split4{u32x2{p01[0], p23[0]}, u32x2{p01[1], p23[1]}};
}

// v_mfma_f32_16x16x16_bf16: A is 16x16 (lane l: row l%16, k-slots 4 (l/16) ..
// + 3), B is 16x16 (lane l: k-slots 4 (l/16) .. + 3, column l%16), C as
// v_mfma_f32_16x16x4f32's; operands given as two words of BF16 pairs.
__device__ __forceinline__ flt4 mfma_bf16_k16(u32x2 a, u32x2 b, flt4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(sh4, a), __builtin_bit_cast(sh4, b), c, 0,
                                                     0, 0);
}

// a b in BF16x3: the three products in the order of the ladder's split kernels
__device__ __forceinline__ flt4 mfma_bf16x3_k16(const Split4 &a, const Split4 &b, flt4 c) {
    c = mfma_bf16_k16(a.l, b.h, c);
    c = mfma_bf16_k16(a.h, b.l, c);
    return mfma_bf16_k16(a.h, b.h, c);
}

// Per-type view of the 16x16 MFMA tile: accumulator type and the output row
// that register q of lane l holds.
template <typename CT> struct Mma;
template <> struct Mma<double> {
    typedef dbl4 acc;
    static __device__ __forceinline__ int row(int q, int lane) { return 4 * q + (lane >> 4); }
};
template <> struct Mma<float> {
    typedef flt4 acc;
    static __device__ __forceinline__ int row(int q, int lane) { return 4 * (lane >> 4) + q; }
};
