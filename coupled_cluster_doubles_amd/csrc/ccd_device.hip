// MI355X (gfx950) implementation of the CCD amplitude iteration: device
// context, setup tables and the C ABI of include/ccd_amd.h.
//
// Replaces, for one GPU, the hot path of justinlietz/coupled_cluster_doubles:
//   calcT2Diff_pphh   src/calc.cpp:132-231   -> k_ladder_x + k_ladder_v
//   calcT2Diff_phhp   src/calc.cpp:264-339   -> k_ring
//   calcT2Diff_h2pp_h1, calcT2Diff_p2_p1hh   -> chi, psi from k_reduce, applied in k_update
//   8 x mapFromDistributed, 5 x mapToDistributed (src/Chain.hpp:710-802)
//                                            -> k_transpose (2 per iteration) + in-row swaps
//   calcT2_pphh       src/calc.cpp:368-459   -> k_update + k_reduce (which ends with the energy)
//   MBPT2Corr         src/MBPT2Corr.cpp:6-38 -> k_mbpt2 + k_sum_partials
//   loop              src/main.cpp:491-682   -> ccd_solve
//
// HBM layout.  Amplitude-shaped arrays T (ping-pong), G (the ring output in
// the pphh layout) and V^hhpp (as "vt", pphh layout) use the reference pphh
// Chain layout; T^ph and R the phhp layout.  V^pppp is stored for channels
// with holes only, on pairs a<b (a quarter of the ordered block); the ladder
// output D, the antisymmetrised amplitudes and V^hhpp rows are kept on the
// same packed rows.  The one-body intermediates are diagonal (chi_i, psi_a),
// so the reference's 3-body chains and their four remaps reduce to sums of
// W = V o T and in-channel row swaps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <limits>
#include <climits>
#include <mutex>
#include <queue>
#include <utility>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "ccd_amd.h"
#include "device_common.hpp"
#include "internal.hpp"
#include "kernels_ladder.hpp"
#include "ladder_launch.hpp"
#include "kernels_ring.hpp"
#include "kernels_setup.hpp"
#include "kernels_update.hpp"

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

// ------------------------------------------------- scratch of ccd_mbpt2_noalloc
//
// ccd_mbpt2_noalloc keeps its per-tile partial sums in this file-scope array
// instead of context memory, so that the call allocates nothing.  Each GPU
// holds its own copy for the life of the process; in host code `scratch` is
// only a stand-in for it, so kernels get the device address that
// hipGetSymbolAddress returns for the device, looked up once per device when
// a context is created there.  Calls on one GPU share the array: calls that
// overlap on one GPU race.
constexpr int kNoAllocPartials = 1 << 16;   // 512 KB per GPU
static __device__ __align__(256) double scratch[kNoAllocPartials];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

// With `device` current: its address of `scratch` (once per device).
static int init_scratch(int device) {
    if (device < 0 || device >= kMaxDevices)
        return ccd::fail(CCD_E_INVALID, "device %d is beyond the %d devices this library supports", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return CCD_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return ccd::fail(CCD_E_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                         hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double *>(p);
    return CCD_OK;
}

static double *scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

// ------------------------------------------------------------------ context

constexpr int kStale = -1;   // tph_prec / tas_prec: not derived from the current amplitudes, rebuild before iterating
constexpr int kAnyPrec = -2;  // ob_prec: chi = psi = 0 (T = 0), right for either precision

struct ccd_ctx {
    int device = 0;
    int precision = CCD_PREC_FP64;
    bool matrix_free = false;    // V^pppp generated inside the ladder, never stored
    int lsrc = LV_STORED;        // LadderSource of the pppp ladder
    int ncu = 1;                 // compute units of the device
    ccd_ladder_plan plan{};      // launch plan of the pppp ladder
    LadderGen lgen{};
    hipStream_t stream = nullptr;         // main stream: every entry point is ordered on it
    hipStream_t aux[3] = {nullptr, nullptr, nullptr};  // concurrent phases of an iteration
    hipEvent_t sync_ev[8] = {};
    int nsp = 0, nh = 0, np = 0;
    int64_t nt2 = 0;
    int64_t bytes = 0;
    std::vector<DChan> chans;
    std::vector<DMod> mods;
    std::vector<DTile> rtiles, ltiles, utiles, xtiles, ttiles;  // packed-row, ladder, update, X, T X tiles
    std::vector<DTile> xftiles;  // X tiles of channels with more than one k segment (k_fold_x)
    int lt_off[NTMAX + 1] = {0};  // ladder tiles grouped by 16-column tile count: [lt_off[n-1], lt_off[n])
    std::vector<DTile> ltiles0;   // the same tiles, channel order inside each class
    // blocks of the ladder launch of class n: lb_grid[n] blocks, block b runs the
    // class's tiles lbptr[lb_base[n] + b] .. lbptr[lb_base[n] + b + 1] - 1
    // (lb_base[n] reserves one entry per tile of the class and one more)
    int lb_base[NTMAX + 1] = {0}, lb_grid[NTMAX + 1] = {0};
    std::vector<int32_t> lbptr;
    int32_t *d_lbptr = nullptr;
    int64_t n_cp = 0;            // column-partial slots
    std::vector<RingItem> ritems;  // ring work items, largest channels first
    int n_ring_wide = 0;           // leading items of channels with more than 5 tiles
    std::vector<RingGroup> rgroups;  // staged-ring blocks: 1-4 tile channels, then 5 tile ones
    int n_rg_five = 0;               // leading groups: the 5-tile channels
    size_t rg_lds[2] = {0, 0};       // LDS of the narrow launch, FP64 / FP32
    int n_tt_narrow = 0;           // leading T X tiles of channels with hh <= 80
    size_t tr_lds = 0;
    int iter_prec = CCD_PREC_FP64;  // arithmetic of the contractions of the next iterations
    // element type T^ph and the antisymmetrised rows (d_tas) were last written
    // in (the ring and the ladder read them in iter_prec); kStale: not derived
    // from the current amplitudes (an entry point computed a caller's term, or
    // a device loop stopped while the last iteration's T remap ran)
    int tph_prec = CCD_PREC_FP64;
    int tas_prec = CCD_PREC_FP64;
    // chi / psi may not be those of the current amplitudes: the psi blocks of
    // the reduction that stopped a device loop may have returned at entry
    bool onebody_stale = false;
    // the V precision of the W sums behind chi and psi (FP32 iterations read
    // V's FP32 rounding; kAnyPrec: chi = psi = 0, from T = 0)
    int ob_prec = CCD_PREC_FP64;
    // device tables
    DChan *d_chans = nullptr;
    DMod *d_mods = nullptr;
    DTile *d_rtiles = nullptr, *d_ltiles = nullptr, *d_utiles = nullptr, *d_xtiles = nullptr, *d_ttiles = nullptr;
    DTile *d_xftiles = nullptr;
    int2 *d_ppl = nullptr, *d_hhl = nullptr, *d_phl = nullptr, *d_hpl = nullptr;
    int32_t *d_hh_old = nullptr, *d_prow = nullptr, *d_prow_sw = nullptr;   // d_hh_old: caller column of each internal hh column
    int32_t *d_red_ptr = nullptr, *d_red_idx = nullptr;   // reduction lists: ranges, places in d_red
    int64_t n_red = 0;
    int32_t *d_tr_base = nullptr, *d_addr_pp = nullptr, *d_addr_ph = nullptr;
    // internal row order of the pphh layout (see build): element maps to and
    // from the caller's order, and the caller's row of each internal pp row
    int32_t *d_t2_in = nullptr, *d_t2_out = nullptr, *d_pp_old = nullptr;
    uint16_t *d_perm_ph = nullptr, *d_perm_pp = nullptr;
    // device state and interaction
    double *d_eps = nullptr, *d_onebody = nullptr;  // chi [nh] then psi [np]
    double *d_t[2] = {nullptr, nullptr};
    int cur = 0;
    double *d_dp = nullptr, *d_tph = nullptr, *d_r = nullptr, *d_g = nullptr, *d_x = nullptr;
    float *d_dsym = nullptr;     // class 1's symmetric partial sums (LK_BF16X3SYM; FP32 contexts)
    double *d_xp = nullptr;      // k-segment partial blocks of X (channels of more than one segment)
    double *d_tas = nullptr, *d_red = nullptr;  // d_red: W partial sums (column partials, then row sums)
    double *d_vt = nullptr, *d_vtp = nullptr, *d_vh = nullptr, *d_vp = nullptr, *d_va = nullptr,
           *d_vb = nullptr;
    double *d_partials = nullptr, *d_scalar = nullptr;
    double *d_zero = nullptr;    // kZeroDoubles zeros, never written: the target of masked-off loads
    LoopCtl *d_ctl = nullptr;    // device loop state
    int *d_ticket = nullptr;     // last-block counter of k_reduce (zero between launches)
    LoopCtl *h_ctl = nullptr;    // pinned staging for it
    LoopCtl *h_ctl2 = nullptr;   // pinned: the loop state after the batch in flight (run_loop)
    hipEvent_t ctl_ev[2] = {};   // the copies of the loop state into h_ctl / h_ctl2 have landed
    // schedule of the untimed iterations (CCD_SCHED_*, never AUTO once tuned)
    // and the tuner's measurements, microseconds per iteration per schedule
    int sched = CCD_SCHED_STREAMS;
    double sched_us[3] = {0, 0, 0};
    // kSolveBatch iterations captured as a graph, by the parity of cur at its
    // start; the capture holds the kernels of one precision and mixing
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int g_prec[2] = {-1, -1}, g_sched[2] = {-1, -1};
    double g_mixing[2] = {0, 0};
    double *d_elog = nullptr;    // energy of every iteration of a device-run loop (kElogCap)
    double *h_scalar = nullptr;  // pinned
    int32_t *d_qn = nullptr;
    double *d_table = nullptr;
    int2 *d_rec1 = nullptr, *d_rec2 = nullptr, *d_rec1f = nullptr, *d_rec2f = nullptr, *d_rec1w = nullptr;  // generated-V pair records
    RingItem *d_ritems = nullptr;
    RingGroup *d_rgroups = nullptr;
    float *d_vp32 = nullptr, *d_va32 = nullptr, *d_vb32 = nullptr, *d_vt32 = nullptr;  // FP32 copies (mixed precision)
    double *d_gtab = nullptr;                   // generated-V electron gas table
    int64_t n_vp = 0, n_tp = 0, n_dp = 0, n_hhhh = 0, n_xp = 0, n_b = 0, n_sym = 0;
    double flops = 0, flops_pppp = 0;
    // a batch of independent solves merged into this context (nprob = 1 otherwise)
    int nprob = 1;
    std::vector<int32_t> hlo;    // holes of solve p: [hlo[p], hlo[p+1])
    std::vector<int64_t> t2_lo;  // its amplitudes in the pphh layout: [t2_lo[p], t2_lo[p+1])
    std::vector<int32_t> ulo;    // its update tiles: [ulo[p], ulo[p+1])
    int32_t *d_hlo = nullptr;
    int32_t *d_pprob = nullptr;  // solve of each particle (transposition blocks)
    int2 *d_ptab = nullptr;      // per-solve gas tables of a matrix-free batch (LadderGen::ptab)
    LoopCtl *d_pctl = nullptr;   // loop state per solve (== d_ctl for a single solve)
    LoopCtl *h_pctl = nullptr;   // pinned staging for it (batches only)
    GenPart *d_gpart = nullptr;
    std::vector<void *> allocs;
};

// How a merged problem splits into solves (one solve: everything).
struct BatchInfo {
    int nprob = 1;
    std::vector<int32_t> chan_lo;   // direct (and mod) channels of solve p: [chan_lo[p], chan_lo[p+1])
    std::vector<int32_t> hlo, plo;  // its holes / particles (merged state numbers)
    std::vector<GenPart> parts;     // generator parameters per solve (gen_kind kGenBatch)
    int64_t table_len = 0;          // entries of the merged direct-term table
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
    for (hipStream_t s : c->aux)
        if (s) (void)hipStreamSynchronize(s);
    for (hipGraphExec_t g : c->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    for (hipEvent_t e : c->sync_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->ctl_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->h_ctl2) (void)hipHostFree(c->h_ctl2);
    for (hipStream_t s : c->aux)
        if (s) (void)hipStreamDestroy(s);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->h_scalar) (void)hipHostFree(c->h_scalar);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    if (c->h_pctl) (void)hipHostFree(c->h_pctl);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ccd::fail(CCD_E_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return CCD_OK;
}

// The per-solve loop states the contraction kernels of a batch read to skip
// stopped solves (nullptr for one problem: its kernels return at entry once
// its loop has stopped).
const LoopCtl *skip_ctl(const ccd_ctx *c) { return c->nprob > 1 ? c->d_pctl : nullptr; }

UpdArgs upd_args(const ccd_ctx *c, const double *t_old, double *t_new, double mixing) {
    UpdArgs a{};
    a.tiles = c->d_utiles; a.chans = c->d_chans; a.ppl = c->d_ppl; a.hhl = c->d_hhl;
    a.prow = c->d_prow; a.prow_sw = c->d_prow_sw;
    a.eps = c->d_eps; a.chi = c->d_onebody; a.psi = c->d_onebody + c->nh;
    a.vt = c->d_vt; a.vt32 = c->d_vt32; a.dp = c->d_dp; a.g = c->d_g; a.t_old = t_old; a.t_new = t_new;
    a.tas = c->d_tas; a.red = c->d_red; a.n_cp = c->n_cp; a.mixing = mixing; a.nh = c->nh; a.ctl = c->d_ctl;
    a.pctl = c->d_pctl;
    a.dsym = c->iter_prec == CCD_PREC_FP32 && c->plan.kind[CCD_PREC_FP32][0] == CCD_LK_BF16X3SYM ? c->d_dsym : nullptr;
    return a;
}

// --- launch helpers, one per phase of the iteration; each takes the stream
// it launches on

template <typename CT, bool WIDE>
void launch_ring_ct(ccd_ctx *c, const CT *va, const CT *vb, int lo, int hi, hipStream_t st) {
    const int n = hi - lo;
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + kThreads / 64 - 1) / (kThreads / 64)));
    hipLaunchKernelGGL((k_ring<CT, WIDE>), grid, dim3(kThreads), 0, st, c->d_ritems + lo, n, c->d_mods,
                       reinterpret_cast<const CT *>(c->d_tph), va, vb, c->d_zero, reinterpret_cast<CT *>(c->d_r), c->d_ctl,
                       skip_ctl(c));
}

template <typename CT>
size_t ring_stage_bytes(int nt) {
    switch (nt) {
        case 1: return RingStage<CT, 1>::BYTES;
        case 2: return RingStage<CT, 2>::BYTES;
        case 3: return RingStage<CT, 3>::BYTES;
        case 4: return RingStage<CT, 4>::BYTES;
        default: return RingStage<CT, 5>::BYTES;
    }
}

// LDS-staged ring: one launch over the groups of 1-4 tile channels, one over
// the 5-tile channels (whose register and LDS footprint would lower the
// occupancy of the others).  The 5-tile blocks have a wave per column tile
// for FP64 and BF16x3 products (2.5-3 us less of a 40-100 us ring at 66
// electrons) and four waves for plain FP32 products, which measured 3 us
// slower with five.
template <typename CT, bool SPLIT>
constexpr int ring_five_waves() { return sizeof(CT) == 8 || SPLIT ? 5 : 4; }

template <typename CT, bool SPLIT = false>
void launch_ring_staged(ccd_ctx *c, const CT *va, const CT *vb, hipStream_t st) {
    const int n5 = c->n_rg_five, n = (int)c->rgroups.size();
    if (n5 > 0) {
        constexpr int NW = ring_five_waves<CT, SPLIT>();
        const size_t lds = RingStage<CT, 5>::BYTES;
        hipLaunchKernelGGL((k_ring_staged<CT, 5, 5, SPLIT, NW>), dim3((unsigned)n5), dim3(NW * 64), lds, st,
                           c->d_rgroups, c->d_mods, reinterpret_cast<const CT *>(c->d_tph), va, vb, c->d_zero,
                           reinterpret_cast<CT *>(c->d_r), c->d_ctl, skip_ctl(c));
    }
    if (n > n5) {
        const size_t lds = c->rg_lds[sizeof(CT) == 8 ? 0 : 1];
        hipLaunchKernelGGL((k_ring_staged<CT, 1, 4, SPLIT>), dim3((unsigned)(n - n5)), dim3(kThreads), lds, st,
                           c->d_rgroups + n5, c->d_mods, reinterpret_cast<const CT *>(c->d_tph), va, vb, c->d_zero,
                           reinterpret_cast<CT *>(c->d_r), c->d_ctl, skip_ctl(c));
    }
}

// The staged instances that may use more than 64 KB of LDS need the limit raised.
template <typename CT, int LO, int HI, bool SPLIT = false, int NW = 4>
int allow_ring_lds(size_t bytes) {
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ring_staged<CT, LO, HI, SPLIT, NW>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return CCD_OK;
}

int allow_ring_lds_all() {
    int rc;
    if ((rc = allow_ring_lds<double, 1, 4>(RingStage<double, 4>::BYTES)) ||
        (rc = allow_ring_lds<double, 5, 5, false, ring_five_waves<double, false>()>(RingStage<double, 5>::BYTES)) ||
        (rc = allow_ring_lds<float, 5, 5, false, ring_five_waves<float, false>()>(RingStage<float, 5>::BYTES)) ||
        (rc = allow_ring_lds<float, 5, 5, true, ring_five_waves<float, true>()>(RingStage<float, 5>::BYTES)))
        return rc;
    return CCD_OK;
}

int launch_ring(ccd_ctx *c, hipStream_t st) {
    const int nw = c->n_ring_wide;
    if (c->iter_prec == CCD_PREC_FP32) {
        launch_ring_ct<float, true>(c, c->d_va32, c->d_vb32, 0, nw, st);
        if (c->plan.ring_fp32 == CCD_RK_BF16X3) launch_ring_staged<float, true>(c, c->d_va32, c->d_vb32, st);
        else launch_ring_staged<float>(c, c->d_va32, c->d_vb32, st);
    } else {
        launch_ring_ct<double, true>(c, c->d_va, c->d_vb, 0, nw, st);
        launch_ring_staged<double>(c, c->d_va, c->d_vb, st);
    }
    return check_launch("k_ring");
}

// pphh -> phhp: amplitudes to T^ph in FP64 (caller arrays of the per-diagram entry points)
int launch_transpose(ccd_ctx *c, bool to_ph, const double *src, double *dst, hipStream_t st) {
    if (c->np == 0 || c->nt2 == 0) return CCD_OK;
    if (to_ph)
        hipLaunchKernelGGL(k_transpose<double>, dim3((unsigned)c->np), dim3(kThreads), c->tr_lds, st, c->d_tr_base,
                           c->d_addr_pp, c->d_addr_ph, c->d_perm_ph, src, dst, c->d_ctl, nullptr, nullptr);
    else
        hipLaunchKernelGGL(k_transpose<double>, dim3((unsigned)c->np), dim3(kThreads), c->tr_lds, st, c->d_tr_base,
                           c->d_addr_ph, c->d_addr_pp, c->d_perm_pp, src, dst, c->d_ctl, nullptr, nullptr);
    return check_launch("k_transpose");
}

// The T remap of an iteration: amplitudes t to the context's T^ph, in the
// arithmetic type of the contractions (FP32 holds the rounding the FP32 ring
// applies anyway)
// force: run even when the loop has stopped (the zero buffer stands in for a
// loop state that is not done)
int launch_tremap(ccd_ctx *c, const double *t, hipStream_t st, bool force = false) {
    c->tph_prec = c->iter_prec;
    if (c->np == 0 || c->nt2 == 0) return CCD_OK;
    const LoopCtl *ctl = force ? reinterpret_cast<const LoopCtl *>(c->d_zero) : c->d_ctl;
    const LoopCtl *pctl = force ? nullptr : skip_ctl(c);
    if (c->iter_prec == CCD_PREC_FP32)
        hipLaunchKernelGGL((k_transpose<double, float>), dim3((unsigned)c->np), dim3(kThreads), c->tr_lds, st,
                           c->d_tr_base, c->d_addr_pp, c->d_addr_ph, c->d_perm_ph, t,
                           reinterpret_cast<float *>(c->d_tph), ctl, c->d_pprob, pctl);
    else
        hipLaunchKernelGGL(k_transpose<double>, dim3((unsigned)c->np), dim3(kThreads), c->tr_lds, st, c->d_tr_base,
                           c->d_addr_pp, c->d_addr_ph, c->d_perm_ph, t, c->d_tph, ctl, c->d_pprob, pctl);
    return check_launch("k_transpose");
}

int launch_antisym(ccd_ctx *c, const double *t);

int launch_reduce(ccd_ctx *c, hipStream_t st, bool bookkeep = true);
template <bool UPDATE>
int launch_update(ccd_ctx *c, const double *t_old, double *t_new, double mixing, hipStream_t st);

// Rebuild what the next iteration reads besides the amplitudes when it is
// not derived from them in the type the iterations read: the antisymmetrised
// rows (after a change of precision), chi and psi (after a device loop
// stopped), T^ph (after either, or after an entry point held a caller's term).
// chi and psi come from the W sums of the current amplitudes, the values the
// update wrote for them.
int sync_state(ccd_ctx *c) {
    int rc;
    double *t = c->d_t[c->cur];
    if (c->tas_prec != c->iter_prec && (rc = launch_antisym(c, t))) return rc;
    if (c->onebody_stale || (c->ob_prec != kAnyPrec && c->ob_prec != c->iter_prec)) {
        if ((rc = launch_update<false>(c, t, nullptr, 0.0, c->stream)) || (rc = launch_reduce(c, c->stream, false)))
            return rc;
        c->onebody_stale = false;
    }
    if (c->tph_prec == c->iter_prec) return CCD_OK;
    return launch_tremap(c, t, c->stream, true);
}

// phhp -> pphh: the ring output R to G, in the contractions' arithmetic type
int launch_remap_r(ccd_ctx *c, hipStream_t st) {
    if (c->np == 0 || c->nt2 == 0) return CCD_OK;
    if (c->iter_prec == CCD_PREC_FP32)
        hipLaunchKernelGGL(k_transpose<float>, dim3((unsigned)c->np), dim3(kThreads), c->tr_lds, st, c->d_tr_base,
                           c->d_addr_ph, c->d_addr_pp, c->d_perm_pp, reinterpret_cast<const float *>(c->d_r),
                           reinterpret_cast<float *>(c->d_g), c->d_ctl, c->d_pprob, skip_ctl(c));
    else
        hipLaunchKernelGGL(k_transpose<double>, dim3((unsigned)c->np), dim3(kThreads), c->tr_lds, st, c->d_tr_base,
                           c->d_addr_ph, c->d_addr_pp, c->d_perm_pp, c->d_r, c->d_g, c->d_ctl, c->d_pprob,
                           skip_ctl(c));
    return check_launch("k_transpose");
}

// The antisymmetrised packed rows of t, in the arithmetic type of the ladder
int launch_antisym(ccd_ctx *c, const double *t) {
    c->tas_prec = c->iter_prec;
    if (c->rtiles.empty()) return CCD_OK;
    if (c->iter_prec == CCD_PREC_FP32)
        hipLaunchKernelGGL(k_antisym<float>, dim3((unsigned)c->rtiles.size()), dim3(kThreads), 0, c->stream,
                           c->d_rtiles, c->d_chans, c->d_prow, c->d_prow_sw, t, reinterpret_cast<float *>(c->d_tas));
    else
        hipLaunchKernelGGL(k_antisym<double>, dim3((unsigned)c->rtiles.size()), dim3(kThreads), 0, c->stream,
                           c->d_rtiles, c->d_chans, c->d_prow, c->d_prow_sw, t, c->d_tas);
    return check_launch("k_antisym");
}

int launch_ladder_x(ccd_ctx *c, hipStream_t st) {
    if (c->xtiles.empty()) return CCD_OK;
    if (c->iter_prec == CCD_PREC_FP32)
        hipLaunchKernelGGL(k_ladder_x<float>, dim3((unsigned)c->xtiles.size()), dim3(kThreads), 0, st, c->d_xtiles,
                           c->d_chans, c->d_vtp, reinterpret_cast<const float *>(c->d_tas), c->d_vh, c->d_zero,
                           c->d_x, c->d_xp, c->d_ctl, skip_ctl(c));
    else
        hipLaunchKernelGGL(k_ladder_x<double>, dim3((unsigned)c->xtiles.size()), dim3(kThreads), 0, st, c->d_xtiles,
                           c->d_chans, c->d_vtp, c->d_tas, c->d_vh, c->d_zero, c->d_x, c->d_xp, c->d_ctl, skip_ctl(c));
    if (!c->xftiles.empty())
        hipLaunchKernelGGL(k_fold_x, dim3((unsigned)c->xftiles.size()), dim3(kThreads), 0, st, c->d_xftiles, c->d_chans,
                           c->d_vh, c->d_xp, c->d_x, c->d_ctl, skip_ctl(c));
    return check_launch("k_ladder_x");
}

// Bytes of the B-tile LDS of class n under kernel kind `kind`.
template <typename CT>
size_t class_bytes(int n, int kind) {
    switch (n) {
        case 1: return ladder_class_bytes<CT, 1>(kind);
        case 2: return ladder_class_bytes<CT, 2>(kind);
        case 3: return ladder_class_bytes<CT, 3>(kind);
        case 4: return ladder_class_bytes<CT, 4>(kind);
        case 5: return ladder_class_bytes<CT, 5>(kind);
        case 6: return ladder_class_bytes<CT, 6>(kind);
        case 7: return ladder_class_bytes<CT, 7>(kind);
        default: return ladder_class_bytes<CT, 8>(kind);
    }
}

// The grouped ladder launch g (classes [LO, HI] = [1, 1], [2, 2], [3, 4],
// [5, 8] or [1, 8] for g = 0..4), running the classes of [lo, hi] among them
// (the others get no blocks): each class its lb_grid blocks and the plan's
// kernel kind; dynamic LDS = the gas table and the B tiles of the largest
// class launched.
template <typename CT>
void launch_grp(ccd_ctx *c, const CT *vp, int g, int lo, int hi, hipStream_t st) {
    static const int glo[5] = {1, 2, 3, 5, 1}, ghi[5] = {1, 2, 4, 8, 8};
    const int LO = glo[g], HI = ghi[g];
    LadderLaunch L{};
    L.dsym = c->d_dsym;
    const int prec = sizeof(CT) == 4 ? CCD_PREC_FP32 : CCD_PREC_FP64;
    int nb = 0;
    size_t need = 0;
    for (int n = LO; n <= HI; n++) {
        L.tile0[n] = c->lt_off[n - 1];
        L.ptr0[n] = c->lb_base[n];
        L.kind[n] = c->plan.kind[prec][n - 1];
        L.blk0[n] = nb;
        if (n >= lo && n <= hi && c->lb_grid[n] > 0) {
            nb += c->lb_grid[n];
            need = std::max(need, class_bytes<CT>(n, L.kind[n]));
        }
    }
    L.blk0[HI + 1] = nb;
    if (nb == 0) return;
    const size_t lds = ladder_tab_bytes(gas_source(c->lsrc) ? c->lgen.ntab : 0) + need;
    const LadderOps<CT> o{c->d_ltiles, c->d_lbptr, c->d_chans, vp, c->lgen, reinterpret_cast<const CT *>(c->d_tas),
                          c->d_zero, reinterpret_cast<CT *>(c->d_dp), c->d_ctl};
    launch_ladder_grp(c->lsrc, g, (unsigned)nb, lds, st, L, o);
}

// Classes [lo, hi] of the pppp ladder in the plan's partition: launches of
// classes 1, 2, 3-4 and 5-8, or one grouped launch.
template <typename CT>
void launch_ladder_v_ct(ccd_ctx *c, const CT *vp, int lo, int hi, hipStream_t st) {
    if (c->plan.partition == CCD_LP_ONE) {
        launch_grp<CT>(c, vp, 4, lo, hi, st);
        return;
    }
    for (int g = 0; g < 4; g++) launch_grp<CT>(c, vp, g, lo, hi, st);
}

int allow_ladder_lds_all() {
    int e;
    if ((e = allow_ladder_lds_f32_gas()) || (e = allow_ladder_lds_f32_other()) || (e = allow_ladder_lds_f64_gas()) ||
        (e = allow_ladder_lds_f64_other()) || (e = allow_ladder_lds_f32_gasb()) || (e = allow_ladder_lds_f64_gasb()))
        return ccd::fail(CCD_E_HIP, "raising the ladder's LDS limit failed: %s", hipGetErrorString((hipError_t)e));
    return CCD_OK;
}

// 1/2 T X into the last D_p slab of every channel
template <typename CT, bool WIDE>
void launch_ladder_tx_ct(ccd_ctx *c, int lo, int hi, hipStream_t st) {
    const int n = hi - lo;
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + kThreads / 64 - 1) / (kThreads / 64)));
    hipLaunchKernelGGL((k_ladder_tx<CT, WIDE>), grid, dim3(kThreads), 0, st, c->d_ttiles + lo, n, c->d_chans,
                       reinterpret_cast<const CT *>(c->d_tas), c->d_x, c->d_zero, reinterpret_cast<CT *>(c->d_dp),
                       c->d_ctl, skip_ctl(c));
}

int launch_ladder_tx(ccd_ctx *c, hipStream_t st) {
    const int n = (int)c->ttiles.size(), nn = c->n_tt_narrow;
    if (c->iter_prec == CCD_PREC_FP32) {
        launch_ladder_tx_ct<float, false>(c, 0, nn, st);
        launch_ladder_tx_ct<float, true>(c, nn, n, st);
    } else {
        launch_ladder_tx_ct<double, false>(c, 0, nn, st);
        launch_ladder_tx_ct<double, true>(c, nn, n, st);
    }
    return check_launch("k_ladder_tx");
}

int launch_ladder_v(ccd_ctx *c, int lo, int hi, hipStream_t st) {
    if (c->ltiles.empty()) return CCD_OK;
    if (c->iter_prec == CCD_PREC_FP32) launch_ladder_v_ct<float>(c, c->d_vp32, lo, hi, st);
    else launch_ladder_v_ct<double>(c, c->d_vp, lo, hi, st);
    return check_launch("k_ladder_v");
}

template <bool UPDATE>
int launch_update(ccd_ctx *c, const double *t_old, double *t_new, double mixing, hipStream_t st) {
    if (UPDATE) c->tas_prec = c->iter_prec;   // the update writes them in OT
    if (c->utiles.empty()) return CCD_OK;
    UpdArgs a = upd_args(c, t_old, t_new, mixing);
    c->ob_prec = c->iter_prec;   // the W sums it writes
    if (c->iter_prec == CCD_PREC_FP32)
        hipLaunchKernelGGL((k_update<UPDATE, float>), dim3((unsigned)c->utiles.size()), dim3(kThreads), 0, st, a);
    else
        hipLaunchKernelGGL((k_update<UPDATE, double>), dim3((unsigned)c->utiles.size()), dim3(kThreads), 0, st, a);
    return check_launch("k_update");
}

// bookkeep = false: chi and psi only, whatever the loop state (no energy, no
// loop step)
int launch_reduce(ccd_ctx *c, hipStream_t st, bool bookkeep) {
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)c->nsp), dim3(kThreads), 0, st, c->d_red_ptr, c->d_red_idx, c->nh,
                       c->d_red,
                       c->d_onebody, c->d_ticket, c->d_scalar, c->d_ctl, c->d_elog, c->nprob, c->d_hlo,
                       c->d_pctl, bookkeep ? 1 : 0);
    return check_launch("k_reduce");
}

// The auxiliary streams start after everything queued so far on the main
// stream (fork) / the main stream continues after everything queued on them
// (join).  Untimed iterations run between a fork and a join.
int fork_aux(ccd_ctx *c) {
    HIP_TRY(hipEventRecord(c->sync_ev[6], c->stream));
    for (hipStream_t s : c->aux) HIP_TRY(hipStreamWaitEvent(s, c->sync_ev[6], 0));
    return CCD_OK;
}

int join_aux(ccd_ctx *c) {
    for (hipStream_t s : c->aux) {
        HIP_TRY(hipEventRecord(c->sync_ev[7], s));
        HIP_TRY(hipStreamWaitEvent(c->stream, c->sync_ev[7], 0));
    }
    return CCD_OK;
}

// One iteration.  Timed (ev = CCD_KT_COUNT + 1 events): every launch group in
// order on the main stream, event k before group k (CCD_KT_* order: ring, R
// remap, X, 1/2 T X, the pppp ladder classes 1, 2, 3-4 and 5-8, update, T
// remap, reductions) and one after the last.  Untimed, one stream
// (CCD_SCHED_SERIAL*): the same launches in dependency order.  Untimed, four
// streams (CCD_SCHED_STREAMS): the independent phases run concurrently,
// ordered by events --
//   main: pppp ladder (1 and 2 tiles)        | update
//   aux1: ring, R remap                      | after the update: T remap (which
//                                            |   the next iteration's ring follows)
//   aux2: X, 1/2 T X                         |
//   aux3: pppp ladder (3+ tiles)             | after the update: reductions +
//                                            |   energy (which the next update waits for)
// (the pppp ladder does not need X; aux2 starts after the previous update).
// Untimed four-stream iterations run between fork_aux and join_aux
// (begin_run, end_run); the loop state is final only after the join.
int run_iteration(ccd_ctx *c, double mixing, hipEvent_t *ev, int reps = 1) {
    const double *t_old = c->d_t[c->cur];
    double *t_new = c->d_t[1 - c->cur];
    int rc;
    if (ev) {
        // each group launched `reps` times back to back (every group but the
        // reductions rewrites the same outputs from the same inputs; repeated
        // reductions count extra iterations of the free run's loop state)
        const hipStream_t st = c->stream;
        int k = 0;
        auto group = [&](auto &&launch) -> int {
            HIP_TRY(hipEventRecord(ev[k++], st));
            for (int r = 0; r < reps; r++)
                if (int rg = launch()) return rg;
            return CCD_OK;
        };
        if ((rc = group([&] { return launch_ring(c, st); })) || (rc = group([&] { return launch_remap_r(c, st); })) ||
            (rc = group([&] { return launch_ladder_x(c, st); })) ||
            (rc = group([&] { return launch_ladder_tx(c, st); })) ||
            (rc = group([&] { return launch_ladder_v(c, 1, 1, st); })) ||
            (rc = group([&] { return launch_ladder_v(c, 2, 2, st); })) ||
            (rc = group([&] { return launch_ladder_v(c, 3, 4, st); })) ||
            (rc = group([&] { return launch_ladder_v(c, 5, NTMAX, st); })) ||
            (rc = group([&] { return launch_update<true>(c, t_old, t_new, mixing, st); })) ||
            (rc = group([&] { return launch_tremap(c, t_new, st); })) ||
            (rc = group([&] { return launch_reduce(c, st); })))
            return rc;
        HIP_TRY(hipEventRecord(ev[k++], st));
        static_assert(CCD_KT_COUNT == 11, "one event per launch group and one at the end");
    } else if (c->sched == CCD_SCHED_SERIAL || c->sched == CCD_SCHED_SERIAL_GRAPH) {
        const hipStream_t st = c->stream;
        if ((rc = launch_ladder_x(c, st)) || (rc = launch_ladder_tx(c, st)) || (rc = launch_ring(c, st)) ||
            (rc = launch_remap_r(c, st)) || (rc = launch_ladder_v(c, 1, NTMAX, st)) ||
            (rc = launch_update<true>(c, t_old, t_new, mixing, st)) || (rc = launch_tremap(c, t_new, st)) ||
            (rc = launch_reduce(c, st)))
            return rc;
    } else {
        const hipStream_t s0 = c->stream, s1 = c->aux[0], s2 = c->aux[1], s3 = c->aux[2];
        hipEvent_t *e = c->sync_ev;
        // e[4] still holds the previous iteration's update (or nothing in the
        // first iteration of a batch, which fork_aux orders instead)
        HIP_TRY(hipStreamWaitEvent(s2, e[4], 0));
        if ((rc = launch_ladder_x(c, s2)) || (rc = launch_ladder_tx(c, s2))) return rc;
        HIP_TRY(hipEventRecord(e[2], s2));
        if ((rc = launch_ring(c, s1)) || (rc = launch_remap_r(c, s1))) return rc;
        HIP_TRY(hipEventRecord(e[0], s1));
        if (c->plan.partition == CCD_LP_ONE) {   // the whole ladder in one launch on the main stream
            if ((rc = launch_ladder_v(c, 1, NTMAX, s0))) return rc;
        } else {
            if ((rc = launch_ladder_v(c, 1, 2, s0)) || (rc = launch_ladder_v(c, 3, NTMAX, s3))) return rc;
        }
        HIP_TRY(hipEventRecord(e[3], s3));
        HIP_TRY(hipStreamWaitEvent(s0, e[0], 0));
        HIP_TRY(hipStreamWaitEvent(s0, e[2], 0));
        HIP_TRY(hipStreamWaitEvent(s0, e[3], 0));
        if ((rc = launch_update<true>(c, t_old, t_new, mixing, s0))) return rc;
        HIP_TRY(hipEventRecord(e[4], s0));
        HIP_TRY(hipStreamWaitEvent(s1, e[4], 0));
        if ((rc = launch_tremap(c, t_new, s1))) return rc;
        HIP_TRY(hipStreamWaitEvent(s3, e[4], 0));
        if ((rc = launch_reduce(c, s3))) return rc;
    }
    c->cur = 1 - c->cur;
    return CCD_OK;
}

// kSolveBatch untimed one-stream iterations captured as one graph, starting
// at the current parity of cur (the capture runs nothing: cur and the
// host-side bookkeeping of the launches end where they started, kSolveBatch
// being even).  Only kernel launches are captured: no event, memset or copy.
int capture_batch(ccd_ctx *c, double mixing, hipGraph_t *g) {
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = CCD_OK;
    for (int k = 0; k < kSolveBatch && !rc; k++) rc = run_iteration(c, mixing, nullptr);
    const hipError_t ee = hipStreamEndCapture(c->stream, g);   // ends the capture whatever happened
    if (!rc && ee != hipSuccess) rc = ccd::fail(CCD_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
    if (rc && *g) {
        (void)hipGraphDestroy(*g);
        *g = nullptr;
    }
    return rc;
}

int drop_graphs(ccd_ctx *c) {
    for (int k = 0; k < 2; k++) {
        if (c->gexec[k]) HIP_TRY(hipGraphExecDestroy(c->gexec[k]));
        c->gexec[k] = nullptr;
        c->g_prec[k] = c->g_sched[k] = -1;
    }
    return CCD_OK;
}

// A run of untimed iterations starts with begin_run (four streams: the aux
// streams start after everything queued on the main stream) and ends with
// end_run (the main stream continues after everything queued on them).
// Inside a run, iterations carry their own cross-stream ordering.
int begin_run(ccd_ctx *c) { return c->sched == CCD_SCHED_STREAMS ? fork_aux(c) : CCD_OK; }
int end_run(ccd_ctx *c) { return c->sched == CCD_SCHED_STREAMS ? join_aux(c) : CCD_OK; }

// The stream whose work writes the loop state last in an iteration (the
// reductions'): a copy of the state queued there follows the iteration.
hipStream_t ctl_stream(const ccd_ctx *c) { return c->sched == CCD_SCHED_STREAMS ? c->aux[2] : c->stream; }

// n untimed iterations inside a run: a graph replay for a whole batch under
// the graph schedule, else launched one by one.
int run_batch(ccd_ctx *c, double mixing, int n) {
    if (n <= 0) return CCD_OK;
    if (c->sched == CCD_SCHED_SERIAL_GRAPH && n == kSolveBatch) {
        const int k = c->cur;
        if (!c->gexec[k] || c->g_prec[k] != c->iter_prec || c->g_sched[k] != c->sched || c->g_mixing[k] != mixing) {
            if (c->gexec[k]) HIP_TRY(hipGraphExecDestroy(c->gexec[k]));
            c->gexec[k] = nullptr;
            hipGraph_t g = nullptr;
            if (int rc = capture_batch(c, mixing, &g)) return rc;
            const hipError_t e = hipGraphInstantiate(&c->gexec[k], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                c->gexec[k] = nullptr;
                return ccd::fail(CCD_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
            }
            c->g_prec[k] = c->iter_prec;
            c->g_sched[k] = c->sched;
            c->g_mixing[k] = mixing;
        }
        HIP_TRY(hipGraphLaunch(c->gexec[k], c->stream));
        c->tph_prec = c->iter_prec;   // what the replayed T remaps and updates write
        c->tas_prec = c->iter_prec;
        return CCD_OK;
    }
    for (int k = 0; k < n; k++)
        if (int rc = run_iteration(c, mixing, nullptr)) return rc;
    return CCD_OK;
}

// n untimed iterations as one run.
int run_n(ccd_ctx *c, double mixing, int n) {
    int rc = begin_run(c);
    for (int k = 0; k < n && !rc; k += kSolveBatch) rc = run_batch(c, mixing, std::min(kSolveBatch, n - k));
    if (!rc) rc = end_run(c);
    return rc;
}

// Caller amplitudes (the reference's pphh order) to an internal-order device
// array, through a device scratch array of nt2 doubles; and back.
int t2_upload(ccd_ctx *c, const double *host, double *dst, double *scratch) {
    HIP_TRY(hipMemcpyAsync(scratch, host, sizeof(double) * c->nt2, hipMemcpyHostToDevice, c->stream));
    if (c->nt2 > 0) {
        const unsigned g = (unsigned)std::min<int64_t>((c->nt2 + kThreads - 1) / kThreads, 8192);
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(kThreads), 0, c->stream, scratch, c->d_t2_in, dst, c->nt2);
    }
    return check_launch("k_gather");
}

int t2_download(ccd_ctx *c, const double *src, double *host, double *scratch) {
    if (c->nt2 > 0) {
        const unsigned g = (unsigned)std::min<int64_t>((c->nt2 + kThreads - 1) / kThreads, 8192);
        hipLaunchKernelGGL(k_gather, dim3(g), dim3(kThreads), 0, c->stream, src, c->d_t2_out, scratch, c->nt2);
    }
    if (int rc = check_launch("k_gather")) return rc;
    HIP_TRY(hipMemcpyAsync(host, scratch, sizeof(double) * c->nt2, hipMemcpyDeviceToHost, c->stream));
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
    } else if (p->gen_kind == CCD_GEN_PAIRING || p->gen_kind == CCD_GEN_ELECTRON_GAS ||
               p->gen_kind == CCD_GEN_INF_MATTER) {
        if (!p->qn) return ccd::fail(CCD_E_INVALID, "generator needs qn");
        if (p->gen_kind == CCD_GEN_ELECTRON_GAS && (!p->gen_table || p->gen_nmax < 0))
            return ccd::fail(CCD_E_INVALID, "electron-gas generator needs gen_table and gen_nmax");
        if (p->gen_kind == CCD_GEN_INF_MATTER && (!p->gen_table || !(p->gen_g > 0)))
            return ccd::fail(CCD_E_INVALID, "infinite-matter generator needs gen_table (3 prefactors) and gen_g = L");
    } else {
        return ccd::fail(CCD_E_INVALID, "unknown gen_kind %d", p->gen_kind);
    }
    return CCD_OK;
}

// One problem as a batch of one.
BatchInfo single_batch(const ccd_problem *p) {
    BatchInfo bi;
    bi.chan_lo = {0, p->n_chan};
    bi.hlo = {0, p->n_holes};
    bi.plo = {p->n_holes, p->n_sp};
    return bi;
}

// Several problems merged into one: states renumbered holes first (the holes
// of solve 0, of solve 1, ..., then the particles in the same order), the
// channel lists and the caller's V chains back to back in solve order (direct
// and mod channels alike), and for generated V the solve's index in qn[5 s + 4]
// with its parameters in BatchInfo::parts.
struct Merged {
    ccd_problem p{};
    BatchInfo bi;
    std::vector<double> eps, v_hhhh, v_hhpp, v_pppp, v_modA, v_modB, table;
    std::vector<int32_t> pp_dim, hh_dim, ph_dim, hp_dim, pp_pairs, hh_pairs, ph_pairs, hp_pairs, qn;
};

int merge_problems(const ccd_problem *const *probs, int n, Merged &m) {
    int nh = 0, nsp = 0, gen = -1;
    for (int q = 0; q < n; q++) {
        int rc = validate(probs[q]);
        if (rc) return rc;
        const bool g = probs[q]->gen_kind != CCD_GEN_NONE;
        if (gen >= 0 && g != (gen == 1))
            return ccd::fail(CCD_E_INVALID, "a batch needs generated V for every problem or for none");
        gen = g ? 1 : 0;
        nh += probs[q]->n_holes;
        nsp += probs[q]->n_sp;
        if ((int64_t)nsp >= INT32_MAX / 8) return ccd::fail(CCD_E_INVALID, "batch too large");
    }
    BatchInfo &bi = m.bi;
    bi.nprob = n;
    bi.chan_lo.assign(1, 0);
    bi.hlo.assign(1, 0);
    bi.plo.assign(1, nh);
    m.eps.assign(nsp, 0.0);
    if (gen) m.qn.assign(5 * (size_t)nsp, 0);
    int h0 = 0, p0 = nh, nchan = 0, nmod = 0;
    for (int q = 0; q < n; q++) {
        const ccd_problem &p = *probs[q];
        const int qh = p.n_holes;
        auto id = [&](int s) { return s < qh ? h0 + s : p0 + (s - qh); };
        for (int s = 0; s < p.n_sp; s++) {
            m.eps[id(s)] = p.eps[s];
            if (gen) {
                for (int f = 0; f < 4; f++) m.qn[5 * (size_t)id(s) + f] = p.qn[5 * (size_t)s + f];
                m.qn[5 * (size_t)id(s) + 4] = q;
            }
        }
        int64_t npp = 0, nhh = 0, nph = 0, nhp = 0, nhhhh = 0, npppp = 0, nt = 0, nmb = 0;
        for (int k = 0; k < p.n_chan; k++) {
            const int64_t P = p.pp_dim[k], H = p.hh_dim[k];
            m.pp_dim.push_back((int32_t)P);
            m.hh_dim.push_back((int32_t)H);
            npp += P; nhh += H; nhhhh += H * H; npppp += P * P; nt += P * H;
        }
        for (int k = 0; k < p.n_mod; k++) {
            const int64_t PH = p.ph_dim[k], HP = p.hp_dim[k];
            m.ph_dim.push_back((int32_t)PH);
            m.hp_dim.push_back((int32_t)HP);
            nph += PH; nhp += HP; nmb += HP * HP;
        }
        auto add_pairs = [&](std::vector<int32_t> &dst, const int32_t *src, int64_t cnt) {
            for (int64_t k = 0; k < 2 * cnt; k++) {
                const int s = src[k];
                if (s < 0 || s >= p.n_sp) return false;
                dst.push_back(id(s));
            }
            return true;
        };
        if (!add_pairs(m.pp_pairs, p.pp_pairs, npp) || !add_pairs(m.hh_pairs, p.hh_pairs, nhh) ||
            !add_pairs(m.ph_pairs, p.ph_pairs, nph) || !add_pairs(m.hp_pairs, p.hp_pairs, nhp))
            return ccd::fail(CCD_E_INVALID, "batch problem %d: pair entry out of range", q);
        if (gen) {
            GenPart gp{p.gen_kind, p.gen_nmax, p.gen_g, (int64_t)m.table.size()};
            if (p.gen_kind == CCD_GEN_ELECTRON_GAS) {
                for (int s = 0; s < p.n_sp; s++)
                    for (int f = 0; f < 3; f++)
                        if (std::abs(p.qn[5 * s + f]) > p.gen_nmax)
                            return ccd::fail(CCD_E_INVALID, "batch problem %d: state %d has |n| > gen_nmax", q, s);
                const size_t w = 4 * (size_t)p.gen_nmax + 1;
                m.table.insert(m.table.end(), p.gen_table, p.gen_table + w * w * w);
            } else if (p.gen_kind == CCD_GEN_INF_MATTER) {
                // qn[5 s + 4] becomes the solve index: one isospin for every
                // state of a pure-neutron-matter problem, so the element is unchanged
                for (int s = 1; s < p.n_sp; s++)
                    if (p.qn[5 * s + 4] != p.qn[4])
                        return ccd::fail(CCD_E_INVALID, "batch problem %d: infinite matter with two isospins", q);
                m.table.insert(m.table.end(), p.gen_table, p.gen_table + 3);
            }
            bi.parts.push_back(gp);
        } else {
            m.v_hhhh.insert(m.v_hhhh.end(), p.v_hhhh, p.v_hhhh + nhhhh);
            m.v_hhpp.insert(m.v_hhpp.end(), p.v_hhpp, p.v_hhpp + nt);
            m.v_pppp.insert(m.v_pppp.end(), p.v_pppp, p.v_pppp + npppp);
            m.v_modA.insert(m.v_modA.end(), p.v_modA, p.v_modA + nt);
            m.v_modB.insert(m.v_modB.end(), p.v_modB, p.v_modB + nmb);
        }
        nchan += p.n_chan;
        nmod += p.n_mod;
        h0 += qh;
        p0 += p.n_sp - qh;
        bi.chan_lo.push_back(nchan);
        bi.hlo.push_back(h0);
        bi.plo.push_back(p0);
    }
    bi.table_len = (int64_t)m.table.size();
    if (m.table.empty()) m.table.push_back(0.0);
    ccd_problem &o = m.p;
    o.n_sp = nsp; o.n_holes = nh; o.eps = m.eps.data();
    o.n_chan = nchan; o.pp_dim = m.pp_dim.data(); o.hh_dim = m.hh_dim.data();
    o.pp_pairs = m.pp_pairs.data(); o.hh_pairs = m.hh_pairs.data();
    o.n_mod = nmod; o.ph_dim = m.ph_dim.data(); o.hp_dim = m.hp_dim.data();
    o.ph_pairs = m.ph_pairs.data(); o.hp_pairs = m.hp_pairs.data();
    if (gen) {
        o.gen_kind = kGenBatch;
        o.qn = m.qn.data();
        o.gen_table = m.table.data();
    } else {
        o.gen_kind = CCD_GEN_NONE;
        o.v_hhhh = m.v_hhhh.data(); o.v_hhpp = m.v_hhpp.data(); o.v_pppp = m.v_pppp.data();
        o.v_modA = m.v_modA.data(); o.v_modB = m.v_modB.data();
    }
    return CCD_OK;
}

// Pair records of the generated-V ladder (LadderGen).  The electron-gas form
// needs, and checks here per solve: two spin values, conserved total momentum
// and spin in every direct channel, and a direct-term table even in each
// component that fits the LDS table (kGasTableMax entries); the channels of
// a solve that fails (and of every non-gas solve of a batch) use the generic
// generator per element.  A batch gets one folded table per gas solve, back
// to back, with (offset, entries) per solve in ptab; the ladder stages the
// table of the solve it works on.
int build_generated_ladder(ccd_ctx *c, const ccd_problem *p, const BatchInfo &bi, const std::vector<int2> &ppl,
                           const std::vector<int32_t> &prow) {
    // kRecPad zero entries past the last pair: the wave-tile ladder's last
    // chunk of a channel reads records beyond it (their amplitude rows are 0)
    const size_t nrec = prow.size() + kRecPad;
    std::vector<int2> rec1(nrec, int2{0, 0}), rec2(nrec, int2{0, 0}), rec1f(nrec, int2{0, 0}),
        rec2f(nrec, int2{0, 0}), rec1w(nrec, int2{0, 0});
    const int nprob = bi.nprob;
    const int32_t *qn = p->qn;
    // per solve: its generator kind, nmax and full table, the spin values
    // of its states, whether it takes the gas form
    std::vector<char> gas(nprob, 0);
    std::vector<int> mq(nprob, 0), s_lo(nprob, 0), s_hi(nprob, 0);
    std::vector<const double *> tq(nprob, nullptr);
    for (int q = 0; q < nprob; q++) {
        const int kind = nprob == 1 ? p->gen_kind : bi.parts[q].kind;
        mq[q] = nprob == 1 ? p->gen_nmax : bi.parts[q].nmax;
        tq[q] = p->gen_table + (nprob == 1 ? 0 : bi.parts[q].tab_off);
        const int D = 2 * mq[q] + 1;
        gas[q] = kind == CCD_GEN_ELECTRON_GAS && (int64_t)D * D * D + 1 <= kGasTableMax;
        if (!gas[q]) continue;
        bool first = true;
        auto spins = [&](int lo, int hi) {
            for (int st = lo; st < hi; st++) {
                const int sz = qn[5 * st + 3];
                if (first) s_lo[q] = s_hi[q] = sz;
                first = false;
                s_lo[q] = std::min(s_lo[q], sz);
                s_hi[q] = std::max(s_hi[q], sz);
            }
        };
        spins(bi.hlo[q], bi.hlo[q + 1]);
        spins(bi.plo[q], bi.plo[q + 1]);
        auto two_spins = [&](int lo, int hi) {
            for (int st = lo; st < hi; st++)
                if (qn[5 * st + 3] != s_lo[q] && qn[5 * st + 3] != s_hi[q]) return false;
            return true;
        };
        gas[q] = two_spins(bi.hlo[q], bi.hlo[q + 1]) && two_spins(bi.plo[q], bi.plo[q + 1]);
    }
    for (const DChan &ch : c->chans) {   // conserved momentum and spin in every channel
        if (!gas[ch.prob]) continue;
        const int2 f = ppl[ch.pp_off];
        int tot[4];
        for (int k = 0; k < 4; k++) tot[k] = qn[5 * f.x + k] + qn[5 * f.y + k];
        for (int u = 1; u < ch.pp && gas[ch.prob]; u++) {
            const int2 e = ppl[ch.pp_off + u];
            for (int k = 0; k < 4; k++)
                if (qn[5 * e.x + k] + qn[5 * e.y + k] != tot[k]) gas[ch.prob] = 0;
        }
    }
    // the folded tables T(|dx|, |dy|, |dz|) of the gas solves, back to back
    std::vector<double> tab;
    std::vector<int2> ptab(nprob, int2{0, 0});
    int ntab_max = 0;
    for (int q = 0; q < nprob; q++) {
        if (!gas[q]) continue;
        const int m = mq[q], D = 2 * m + 1, w = 4 * m + 1;
        const double *T = tq[q];
        auto at = [&](int dx, int dy, int dz) {
            return T[((size_t)(dx + 2 * m) * w + (dy + 2 * m)) * w + (dz + 2 * m)];
        };
        std::vector<double> f((size_t)D * D * D + 1, 0.0);
        for (int x = 0; x < D && gas[q]; x++)
            for (int y = 0; y < D && gas[q]; y++)
                for (int z = 0; z < D && gas[q]; z++) {
                    const double v = at(x, y, z);
                    for (int sg = 1; sg < 8; sg++)
                        if (at(sg & 1 ? -x : x, sg & 2 ? -y : y, sg & 4 ? -z : z) != v) gas[q] = 0;
                    f[((size_t)x * D + y) * D + z] = v;
                }
        if (!gas[q]) continue;
        ptab[q] = int2{(int)tab.size(), (int)f.size()};
        ntab_max = std::max(ntab_max, (int)f.size());
        tab.insert(tab.end(), f.begin(), f.end());
    }
    // Records: the momentum of a state pre-scaled into 16-bit fields,
    // esize D^2 (n_x + m) | esize D (n_y + m) << 16 and esize (n_z + m), so that
    // two v_sad_u16 of two records give the byte offset of T(|dn|) in a table
    // of esize-byte entries (at most 8 D^3 <= 32768 bytes; D and m of the
    // pair's solve).
    // Channels whose pairs all have one particle of each spin (total spin 0):
    // there <ab||cd> = s(ab) s(cd) T(k_up(ab) - k_up(cd)), with k_up the
    // momentum of the pair's up-spin particle and s = +1 when the first
    // particle is the up one, -1 otherwise (exactly one of the two terms
    // survives, and T is even).  Their records: rec1 = k_up (rec1f for the
    // FP32 table); rec2 = (sign bit of s as a float/double sign mask, 0);
    // rec1w = rec1f with s in the high half of its second word, 0 for +1 and
    // the byte size of the solve's FP32 table for -1 (the FP32 tables are
    // [T, -T]: the two fields' distance selects the copy of sign s(ab) s(cd)).
    static_assert(4 * kGasTableMax < 65536, "the sign field holds the FP32 table's byte size");
    auto scaled_up = [&](int q, int s, uint32_t esize) {   // esize: bytes per table entry
        const int m = mq[q], D = 2 * m + 1;
        const int32_t *a = qn + 5 * (size_t)s;
        const uint32_t x = esize * D * D * (a[0] + m), y = esize * D * (a[1] + m), z = esize * (a[2] + m);
        return int2{(int)(x | y << 16), (int)z};
    };
    // Channels whose pairs have two particles of the same spin (or a basis
    // of one spin): both terms survive for every element (spin0 = 2); rec1,
    // rec2 = the momenta of the two states (rec1f, rec2f for the FP32 table).
    // A gas channel with pairs is one of the two kinds.
    for (DChan &ch : c->chans) {
        const int q = ch.prob;
        ch.spin0 = 0;
        if (gas[q] && ch.npk + ch.ndiag > 0) {
            const int2 f = ppl[ch.pp_off + prow[ch.pr_off]];
            ch.spin0 = qn[5 * f.x + 3] != qn[5 * f.y + 3] && ch.ndiag == 0 ? 1 : 2;
        }
        const uint32_t sbytes = gas[q] ? 4u * (uint32_t)ptab[q].y : 0u;   // bytes of the solve's FP32 table
        for (int u = 0; u < ch.npk + ch.ndiag; u++) {
            const int2 ab = ppl[ch.pp_off + prow[ch.pr_off + u]];
            if (ch.spin0 == 1) {
                const bool first_up = qn[5 * ab.x + 3] == s_hi[q];
                rec1[ch.pr_off + u] = scaled_up(q, first_up ? ab.x : ab.y, 8u);
                rec1f[ch.pr_off + u] = scaled_up(q, first_up ? ab.x : ab.y, 4u);   // the FP32 table
                rec1w[ch.pr_off + u] = rec1f[ch.pr_off + u];
                rec1w[ch.pr_off + u].y |= (int)((first_up ? 0u : sbytes) << 16);
                rec2[ch.pr_off + u] = int2{first_up ? 0 : (int)0x80000000u, 0};
            } else if (ch.spin0 == 2) {
                rec1[ch.pr_off + u] = scaled_up(q, ab.x, 8u);
                rec2[ch.pr_off + u] = scaled_up(q, ab.y, 8u);
                rec1f[ch.pr_off + u] = scaled_up(q, ab.x, 4u);
                rec2f[ch.pr_off + u] = scaled_up(q, ab.y, 4u);
            } else {
                rec1[ch.pr_off + u] = int2{ab.x, ab.y};
            }
        }
    }
    int rc;
    if (!c->chans.empty())   // the spin0 flags
        HIP_TRY(hipMemcpy(c->d_chans, c->chans.data(), sizeof(DChan) * c->chans.size(), hipMemcpyHostToDevice));
    if ((rc = upload(c, &c->d_rec1, rec1)) || (rc = upload(c, &c->d_rec2, rec2)) ||
        (rc = upload(c, &c->d_rec1f, rec1f)) || (rc = upload(c, &c->d_rec2f, rec2f)) ||
        (rc = upload(c, &c->d_rec1w, rec1w)))
        return rc;
    c->lgen.rec1 = c->d_rec1;
    c->lgen.rec2 = c->d_rec2;
    c->lgen.rec1f = c->d_rec1f;
    c->lgen.rec2f = c->d_rec2f;
    c->lgen.rec1w = c->d_rec1w;
    c->lgen.pctl = nprob > 1 ? c->d_pctl : nullptr;
    c->lgen.ptab = nullptr;
    if (!tab.empty()) {
        if ((rc = upload(c, &c->d_gtab, tab))) return rc;
        c->lgen.tab = c->d_gtab;
        c->lgen.ntab = ntab_max;
        if (nprob > 1) {
            if ((rc = upload(c, &c->d_ptab, ptab))) return rc;
            c->lgen.ptab = c->d_ptab;
        }
        c->lsrc = nprob > 1 ? LV_GAS_BATCH : LV_GAS;
    } else {
        c->lsrc = LV_GENERIC;
    }
    return CCD_OK;
}

// The plan a context starts with: the default kernels, 3 persistent blocks
// per compute unit for a generated V (each stages the generator table once;
// class 1 at N=54 37.3 us against 43.6 us with 4) or one block per tile for
// a stored one, the CU-load model's assignment, launches per class group.
void default_plan(ccd_ctx *c) {
    ccd_ladder_plan &p = c->plan;
    p = ccd_ladder_plan{};
    for (int n = 0; n < NTMAX; n++) {
        p.kind[CCD_PREC_FP64][n] = p.kind[CCD_PREC_FP32][n] = CCD_LK_DEFAULT;
        p.blocks_per_cu[n] = c->matrix_free ? 3 : 0;
        p.assign[n] = CCD_LA_MODEL;
    }
    p.partition = CCD_LP_SPLIT;
    p.ring_fp32 = CCD_RK_FP32;
}

// Block runs of class nt's ladder tiles under the plan (host tables ltiles,
// lbptr, lb_grid).  The grid is one block per tile or blocks_per_cu blocks
// per compute unit, block b running a contiguous run of the class's tiles.
// Two assignments are costed with a model of the CU loads (block b on CU
// b mod ncu, cost of a tile = its k chunks + 1): round robin in channel
// order (block b takes tiles b, b+G, ...), and longest tile first to the
// least loaded block.  CCD_LA_MODEL takes the second when its most loaded CU
// carries at least 10% less (the large electron gases) and keeps the channel
// order otherwise (it measured faster where the two are close).  Tiles write
// disjoint outputs: the assignment changes no result.
void plan_class_tiles(ccd_ctx *c, int nt) {
    const int lo = c->lt_off[nt - 1], n = c->lt_off[nt] - lo, ncu = c->ncu;
    const int bpc = c->plan.blocks_per_cu[nt - 1];
    const int G = bpc > 0 ? std::min(n, bpc * ncu) : n;
    c->lb_grid[nt] = G;
    int32_t *ptr = c->lbptr.data() + c->lb_base[nt];
    ptr[0] = 0;
    if (G <= 0) return;
    const int kch = nt <= 5 ? LadderShape<1>::KCH : LadderShape<8>::KCH;
    const int kseg = nt <= 2 ? LadderShape<1>::KSEG : LadderShape<3>::KSEG;
    // the symmetric class-1 kernel: a segment-0 tile covers k from its first
    // row to the end of the channel, a later segment's tile only clears its slab
    const bool sym = nt == 1 && c->plan.kind[CCD_PREC_FP32][0] == CCD_LK_BF16X3SYM;
    std::vector<int64_t> cost(n);
    for (int i = 0; i < n; i++) {   // k chunks of the tile, plus one for its prologue
        const DTile &t = c->ltiles0[lo + i];
        const int npk = c->chans[t.chan].npk;
        const int klen = sym ? (t.aux == 0 ? npk - t.row0 : 0) : std::min(npk, (t.aux + 1) * kseg) - t.aux * kseg;
        cost[i] = (klen + kch - 1) / kch + 1;
    }
    auto cu_max = [&](const std::vector<std::vector<int>> &own) {
        std::vector<int64_t> cu(ncu, 0);
        for (int b = 0; b < G; b++)
            for (int i : own[b]) cu[b % ncu] += cost[i];
        return *std::max_element(cu.begin(), cu.end());
    };
    std::vector<std::vector<int>> rr(G), lpt(G);
    for (int i = 0; i < n; i++) rr[i % G].push_back(i);
    int mode = c->plan.assign[nt - 1];
    // a matrix-free batch: tiles of one solve share its gas table, which a
    // block stages when it moves on to another solve
    if (mode == CCD_LA_MODEL && c->nprob > 1 && c->matrix_free) mode = CCD_LA_RUNS;
    if (mode == CCD_LA_RUNS) {   // contiguous runs of about equal cost, in channel order
        int64_t total = 0, acc = 0;
        for (int i = 0; i < n; i++) total += cost[i];
        std::vector<std::vector<int>> runs(G);
        for (int i = 0; i < n; i++) {
            runs[(size_t)std::min<int64_t>(G - 1, acc * G / std::max<int64_t>(total, 1))].push_back(i);
            acc += cost[i];
        }
        int k = 0;
        for (int b = 0; b < G; b++) {
            for (int i : runs[b]) c->ltiles[lo + k++] = c->ltiles0[lo + i];
            ptr[b + 1] = k;
        }
        return;
    }
    if (mode != CCD_LA_CHANNEL) {
        // XCD groups of the blocks: the dispatcher sends block b to XCD
        // (b + c) mod 8, so the blocks b = x mod 8 share an L2 (CCD_LA_XCD);
        // one group of all blocks otherwise
        const int nx = mode == CCD_LA_XCD ? std::min(G, 8) : 1;
        std::vector<int> grp(n, 0);
        if (nx > 1) {
            // units: the tiles of a channel in order, cut where a unit would
            // pass 1/32 of the class; largest unit first to the least loaded XCD
            int64_t total = 0;
            for (int i = 0; i < n; i++) total += cost[i];
            const int64_t cap = std::max<int64_t>(1, total / 32);
            std::vector<std::pair<int64_t, int>> units;   // (cost, first tile)
            std::vector<int> unit_of(n);
            for (int i = 0; i < n; i++) {
                if (i == 0 || c->ltiles0[lo + i].chan != c->ltiles0[lo + i - 1].chan ||
                    units.back().first + cost[i] > cap)
                    units.push_back({0, i});
                units.back().first += cost[i];
                unit_of[i] = (int)units.size() - 1;
            }
            std::vector<int> uidx(units.size()), ux(units.size());
            for (size_t u = 0; u < units.size(); u++) uidx[u] = (int)u;
            std::stable_sort(uidx.begin(), uidx.end(), [&](int a, int b) { return units[a].first > units[b].first; });
            std::vector<int64_t> xload(nx, 0);
            for (int u : uidx) {
                const int x = (int)(std::min_element(xload.begin(), xload.end()) - xload.begin());
                ux[u] = x;
                xload[x] += units[u].first;
            }
            for (int i = 0; i < n; i++) grp[i] = ux[unit_of[i]];
        }
        std::vector<int> idx(n);
        for (int i = 0; i < n; i++) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost[a] > cost[b]; });
        // least loaded block of the tile's group, lowest index on ties
        typedef std::priority_queue<std::pair<int64_t, int>, std::vector<std::pair<int64_t, int>>,
                                    std::greater<std::pair<int64_t, int>>>
            Heap;
        std::vector<Heap> heap(nx);
        for (int b = 0; b < G; b++) heap[b % nx].push({0, b});
        for (int i : idx) {
            Heap &h = heap[grp[i]];
            auto [load, b] = h.top();
            h.pop();
            lpt[b].push_back(i);
            h.push({load + cost[i], b});
        }
    }
    const bool use_lpt = mode == CCD_LA_LONGEST || mode == CCD_LA_XCD ||
                         (mode == CCD_LA_MODEL && 10 * cu_max(lpt) <= 9 * cu_max(rr));
    const auto &own = use_lpt ? lpt : rr;
    int k = 0;
    for (int b = 0; b < G; b++) {
        for (int i : own[b]) c->ltiles[lo + k++] = c->ltiles0[lo + i];
        ptr[b + 1] = k;
    }
}

void plan_ladder_tiles(ccd_ctx *c) {
    for (int nt = 1; nt <= NTMAX; nt++) plan_class_tiles(c, nt);
}

// The host tables of plan_ladder_tiles to the device (after a change of plan).
int upload_ladder_tiles(ccd_ctx *c) {
    if (!c->ltiles.empty())
        HIP_TRY(hipMemcpyAsync(c->d_ltiles, c->ltiles.data(), sizeof(DTile) * c->ltiles.size(),
                               hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_lbptr, c->lbptr.data(), sizeof(int32_t) * c->lbptr.size(), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

// Host tables: channel descriptors, swaps, packed rows, tiles, transposition
// plans and reduction lists.  Every index a kernel dereferences is produced
// and range-checked here.
int build(ccd_ctx *c, const ccd_problem *p, const BatchInfo &bi) {
    const int nsp = p->n_sp, nh = p->n_holes, np = nsp - nh;
    c->nsp = nsp; c->nh = nh; c->np = np;
    const int nprob = bi.nprob;
    c->nprob = nprob;
    c->hlo = bi.hlo;
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
    if (npp >= INT32_MAX || nhh >= INT32_MAX) return ccd::fail(CCD_E_INVALID, "pair lists too long");
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

    // ---- direct channels: swaps, packed rows, row pairs
    std::vector<int32_t> pp_swap(npp, -1), hh_old(nhh);
    std::vector<int32_t> prow, prow_sw;
    int64_t off_t = 0, off_vp = 0, off_tp = 0, off_dp = 0, off_hhhh = 0, off_xp = 0, ipp = 0, ihh = 0;
    int64_t off_sym = 0;   // floats, in whole blocks of 1024
    c->flops = 0; c->flops_pppp = 0;
    c->t2_lo.assign(nprob + 1, 0);
    std::unordered_map<int64_t, int32_t> pos;
    std::vector<int32_t> order, hord, t2_in, t2_out, pp_old(npp);
    std::vector<std::array<int32_t, 4>> pkey;   // sort keys of a channel's pairs a<b
    std::vector<int2> seg;
    for (int64_t k = 0, o = 0; k < p->n_chan; k++)
        for (int q = 0; q < p->pp_dim[k]; q++) pp_old[o++] = q;
    for (int64_t k = 0, o = 0; k < p->n_chan; k++)
        for (int q = 0; q < p->hh_dim[k]; q++) hh_old[o++] = q;
    int prob = 0;
    for (int k = 0; k < p->n_chan; k++) {
        while (prob + 1 < nprob && k >= bi.chan_lo[prob + 1]) c->t2_lo[++prob] = off_t;
        const int P = p->pp_dim[k], H = p->hh_dim[k];
        if (P > 0 && H > 0) {
            if (H > NTMAX * 16) return ccd::fail(CCD_E_INVALID, "channel %d has hh=%d > %d", k, H, NTMAX * 16);
            pos.clear();
            for (int q = 0; q < P; q++) pos[(int64_t)ppl[ipp + q].x * nsp + ppl[ipp + q].y] = q;
            if ((int)pos.size() != P) return ccd::fail(CCD_E_INVALID, "pp channel %d lists a pair twice", k);
            // internal row order: each pair (a,b) with a<b directly followed
            // by its partner (b,a); then the pairs (a,a).  The update reads
            // both rows of a pair, so its loads cover one contiguous run of
            // rows.  With quantum numbers the pairs a<b go in lexicographic
            // order of the momentum (first three numbers) of their up-spin
            // state, or of a when both spins agree, else in the caller's
            // order: the generated ladder's table reads of neighbouring rows
            // then fall on neighbouring entries (the table runs over |dz|
            // fastest) instead of random banks.
            order.clear();
            pkey.clear();
            for (int q = 0; q < P; q++)
                if (ppl[ipp + q].x < ppl[ipp + q].y) {
                    const int a = ppl[ipp + q].x, b = ppl[ipp + q].y;
                    std::array<int32_t, 4> key{0, 0, 0, q};
                    if (p->qn) {
                        const int32_t *qa = p->qn + 5 * (size_t)a, *qb = p->qn + 5 * (size_t)b;
                        const int32_t *qs = (qa[3] != qb[3] && qb[3] > qa[3]) ? qb : qa;
                        key = {qs[0], qs[1], qs[2], q};
                    }
                    pkey.push_back(key);
                }
            std::sort(pkey.begin(), pkey.end());
            for (const auto &key : pkey) {
                const int q = key[3];
                auto it = pos.find((int64_t)ppl[ipp + q].y * nsp + ppl[ipp + q].x);
                if (it == pos.end()) return ccd::fail(CCD_E_INVALID, "pp channel %d lacks a swapped pair", k);
                order.push_back(q);
                order.push_back(it->second);
            }
            for (int q = 0; q < P; q++)
                if (ppl[ipp + q].x == ppl[ipp + q].y) order.push_back(q);
            if ((int)order.size() != P) return ccd::fail(CCD_E_INVALID, "pp channel %d lacks a swapped pair", k);
            // internal column order, likewise: each hole pair (i,j) with i<j
            // directly followed by (j,i), then the pairs (i,i), so that the
            // update finds the partner column (j,i) of column IJ at IJ ^ 1
            pos.clear();
            for (int q = 0; q < H; q++) pos[(int64_t)hhl[ihh + q].x * nsp + hhl[ihh + q].y] = q;
            if ((int)pos.size() != H) return ccd::fail(CCD_E_INVALID, "hh channel %d lists a pair twice", k);
            hord.clear();
            for (int q = 0; q < H; q++)
                if (hhl[ihh + q].x < hhl[ihh + q].y) {
                    auto it = pos.find((int64_t)hhl[ihh + q].y * nsp + hhl[ihh + q].x);
                    if (it == pos.end()) return ccd::fail(CCD_E_INVALID, "hh channel %d lacks a swapped pair", k);
                    hord.push_back(q);
                    hord.push_back(it->second);
                }
            const int hh2 = (int)hord.size();
            for (int q = 0; q < H; q++)
                if (hhl[ihh + q].x == hhl[ihh + q].y) hord.push_back(q);
            if ((int)hord.size() != H) return ccd::fail(CCD_E_INVALID, "hh channel %d lacks a swapped pair", k);
            seg.assign(hhl.begin() + ihh, hhl.begin() + ihh + H);
            for (int q = 0; q < H; q++) {
                hhl[ihh + q] = seg[hord[q]];
                hh_old[ihh + q] = hord[q];
            }
            seg.assign(ppl.begin() + ipp, ppl.begin() + ipp + P);
            t2_in.resize((size_t)(off_t + (int64_t)P * H));
            t2_out.resize((size_t)(off_t + (int64_t)P * H));
            for (int q = 0; q < P; q++) {
                ppl[ipp + q] = seg[order[q]];
                pp_old[ipp + q] = order[q];
                for (int ij = 0; ij < H; ij++) {
                    t2_in[off_t + (int64_t)q * H + ij] = (int32_t)(off_t + (int64_t)order[q] * H + hord[ij]);
                    t2_out[off_t + (int64_t)order[q] * H + hord[ij]] = (int32_t)(off_t + (int64_t)q * H + ij);
                }
            }
            pos.clear();
            for (int q = 0; q < P; q++) pos[(int64_t)ppl[ipp + q].x * nsp + ppl[ipp + q].y] = q;
            for (int q = 0; q < P; q++) {
                auto it = pos.find((int64_t)ppl[ipp + q].y * nsp + ppl[ipp + q].x);
                if (it == pos.end()) return ccd::fail(CCD_E_INVALID, "pp channel %d lacks a swapped pair", k);
                pp_swap[ipp + q] = it->second;
            }
            pos.clear();
            for (int q = 0; q < H; q++) pos[(int64_t)hhl[ihh + q].x * nsp + hhl[ihh + q].y] = q;
            for (int q = 0; q < H; q++) {   // the update computes the partner column from IJ
                auto it = pos.find((int64_t)hhl[ihh + q].y * nsp + hhl[ihh + q].x);
                if (it == pos.end() || it->second != (q < hh2 ? (q ^ 1) : q))
                    return ccd::fail(CCD_E_INVALID, "hh channel %d: unexpected internal column order", k);
            }
            // pair list: packed rows (a<b) first, then rows (a,a)
            const int pr_off = (int)prow.size();
            for (int q = 0; q < P; q++)
                if (ppl[ipp + q].x < ppl[ipp + q].y) {
                    prow.push_back(q);
                    prow_sw.push_back(pp_swap[ipp + q]);
                }
            const int npk = (int)prow.size() - pr_off;
            for (int q = 0; q < P; q++)
                if (ppl[ipp + q].x == ppl[ipp + q].y) {
                    prow.push_back(q);
                    prow_sw.push_back(q);
                }
            const int ndiag = (int)prow.size() - pr_off - npk;
            for (int u = 0; u < npk + ndiag; u++)   // the update computes the rows from u
                if (prow[pr_off + u] != (u < npk ? 2 * u : npk + u) ||
                    prow_sw[pr_off + u] != (u < npk ? 2 * u + 1 : npk + u))
                    return ccd::fail(CCD_E_INVALID, "pp channel %d: unexpected internal row order", k);
            if ((int64_t)npk * H * 8 >= (int64_t)INT32_MAX)   // the ladder's buffer loads use 32-bit offsets
                return ccd::fail(CCD_E_INVALID, "channel %d has %lld packed amplitudes (limit 2^28)", k,
                                 (long long)npk * H);
            const int ldv = (npk + 15) / 16 * 16;
            const int nseg = ladder_segments(npk, (H + 15) / 16);
            const int nxs = std::max(1, (npk + KXSEG - 1) / KXSEG);
            // class 1 on the symmetric kernel: the rows of row tile J hold J
            // slabs of LROWS x hh partial sums
            const int nrt = (npk + LROWS - 1) / LROWS;
            int32_t sym = -1;
            if (H <= 16 && nrt >= 2) {
                if (off_sym / 1024 >= (int64_t)INT32_MAX)
                    return ccd::fail(CCD_E_INVALID, "symmetric ladder sums too large (channel %d)", k);
                sym = (int32_t)(off_sym / 1024);
                off_sym += ((int64_t)nrt * (nrt - 1) / 2 * LROWS * H + 1023) / 1024 * 1024;
            }
            DChan ch{off_t, off_vp, off_tp, off_hhhh, P, H, (int32_t)ipp, (int32_t)ihh, npk, ldv, pr_off, ndiag,
                     off_dp, nseg, prob, 0, hh2, off_xp, nxs, sym};
            if (nxs > 1) off_xp += (int64_t)nxs * H * H;
            c->chans.push_back(ch);
            off_vp += (int64_t)npk * ldv;
            off_tp += (int64_t)npk * H;
            off_dp += (int64_t)(nseg + 1) * npk * H;
            off_hhhh += (int64_t)H * H;
            c->flops += 2.0 * H * H * P + 2.0 * P * H * H + 2.0 * P * P * H;
            c->flops_pppp += 2.0 * P * P * H;
        }
        off_t += (int64_t)P * H;
        ipp += P; ihh += H;
    }
    c->nt2 = off_t;
    while (prob + 1 < nprob) c->t2_lo[++prob] = off_t;
    c->t2_lo[nprob] = off_t;
    c->n_vp = off_vp; c->n_tp = off_tp; c->n_dp = off_dp; c->n_hhhh = off_hhhh; c->n_xp = off_xp;
    c->n_sym = off_sym;
    if (c->nt2 >= (int64_t)INT32_MAX)
        return ccd::fail(CCD_E_INVALID, "T2 has %lld elements (limit 2^31)", (long long)c->nt2);

    // ---- (particle, hole) slots of the position tables: solve q owns a block
    // of np_q x nh_q slots (a pair of two solves has none)
    std::vector<int32_t> sprob(nsp, 0);
    std::vector<int64_t> sbase(nprob + 1, 0);
    for (int q = 0; q < nprob; q++) {
        for (int x = bi.hlo[q]; x < bi.hlo[q + 1]; x++) sprob[x] = q;
        for (int x = bi.plo[q]; x < bi.plo[q + 1]; x++) sprob[x] = q;
        sbase[q + 1] = sbase[q] + (int64_t)(bi.hlo[q + 1] - bi.hlo[q]) * (bi.plo[q + 1] - bi.plo[q]);
    }
    auto slot = [&](int a, int i) -> int64_t {   // particle a, hole i
        const int q = sprob[a];
        if (sprob[i] != q) return -1;
        return sbase[q] + (int64_t)(a - bi.plo[q]) * (bi.hlo[q + 1] - bi.hlo[q]) + (i - bi.hlo[q]);
    };
    const size_t nslot = (size_t)sbase[nprob];

    // ---- mod channels and position tables (host only)
    std::vector<int32_t> colidx(nslot, -1);
    std::vector<int32_t> mod_of_ph(nslot, -1), mod_of_hp(nslot, -1);
    int64_t off_ph = 0, off_b = 0, iph = 0, ihp = 0;
    for (int k = 0; k < p->n_mod; k++) {
        const int PH = p->ph_dim[k], HP = p->hp_dim[k];
        for (int q = 0; q < PH; q++) {
            int a = phl[iph + q].x, i = phl[iph + q].y;
            const int64_t s = slot(a, i);
            if (s < 0) return ccd::fail(CCD_E_INVALID, "ph pair (%d,%d) spans two solves", a, i);
            if (mod_of_ph[s] >= 0) return ccd::fail(CCD_E_INVALID, "ph pair (%d,%d) listed twice", a, i);
            mod_of_ph[s] = k;
        }
        for (int q = 0; q < HP; q++) {
            int j = hpl[ihp + q].x, b = hpl[ihp + q].y;
            const int64_t s = slot(b, j);
            if (s < 0) return ccd::fail(CCD_E_INVALID, "hp pair (%d,%d) spans two solves", j, b);
            if (mod_of_hp[s] >= 0) return ccd::fail(CCD_E_INVALID, "hp pair (%d,%d) listed twice", j, b);
            colidx[s] = q;
            mod_of_hp[s] = k;
        }
        if (PH > 0 && HP > 0) {
            if (PH > NTMAX * 16 || HP > NTMAX * 16)
                return ccd::fail(CCD_E_INVALID, "mod channel %d has ph=%d hp=%d > %d", k, PH, HP, NTMAX * 16);
            for (int jt = 0; jt < (HP + 15) / 16; jt++) c->ritems.push_back(RingItem{(int32_t)c->mods.size(), jt});
            int32_t mprob = 0;   // channels (direct and mod alike) of solve q: [chan_lo[q], chan_lo[q+1])
            while (mprob + 1 < nprob && k >= bi.chan_lo[mprob + 1]) mprob++;
            c->mods.push_back(DMod{off_ph, off_b, PH, HP, (int32_t)iph, (int32_t)ihp, mprob, 0});
            off_b += (int64_t)HP * HP;
            c->flops += 2.0 * HP * HP * PH + 2.0 * PH * HP * HP;
        }
        off_ph += (int64_t)PH * HP;
        iph += PH; ihp += HP;
    }
    c->n_b = off_b;
    if (off_ph != c->nt2)
        return ccd::fail(CCD_E_INVALID, "phhp layout has %lld elements, pphh has %lld", (long long)off_ph,
                         (long long)c->nt2);
    {   // ring items, channels with more 16-wide tiles first (stable: channel order inside a size)
        auto tiles_of = [&](const RingItem &w) {
            return (std::max(c->mods[w.mod].ph, c->mods[w.mod].hp) + 15) / 16;
        };
        std::stable_sort(c->ritems.begin(), c->ritems.end(),
                         [&](const RingItem &a, const RingItem &b) { return tiles_of(a) > tiles_of(b); });
        c->n_ring_wide = 0;
        while (c->n_ring_wide < (int)c->ritems.size() && tiles_of(c->ritems[c->n_ring_wide]) > 5) c->n_ring_wide++;
        // blocks of the LDS-staged kernel: channels of one tile count per
        // block (4 of 1 tile, 2 of 2, else 1), more tiles first
        c->rgroups.clear();
        c->rg_lds[0] = c->rg_lds[1] = 0;
        for (int nt = 5; nt >= 1; nt--) {
            const int per = nt == 1 ? 4 : nt == 2 ? 2 : 1;
            const size_t before = c->rgroups.size();
            RingGroup g{{-1, -1, -1, -1}};
            int fill = 0;
            for (size_t k = 0; k < c->mods.size(); k++) {
                if ((std::max(c->mods[k].ph, c->mods[k].hp) + 15) / 16 != nt) continue;
                g.mod[fill++] = (int32_t)k;
                if (fill == per) {
                    c->rgroups.push_back(g);
                    g = RingGroup{{-1, -1, -1, -1}};
                    fill = 0;
                }
            }
            if (fill) c->rgroups.push_back(g);
            if (nt == 5) c->n_rg_five = (int)c->rgroups.size();   // the 5-tile groups lead
            if (nt <= 4 && c->rgroups.size() > before) {
                c->rg_lds[0] = std::max(c->rg_lds[0], ring_stage_bytes<double>(nt));
                c->rg_lds[1] = std::max(c->rg_lds[1], ring_stage_bytes<float>(nt));
            }
        }
    }

    // ---- per-particle transposition plans between the two layouts:
    // particle a's block holds the elements (a,b,i,j); its pphh slots run over
    // the rows (a,b) in channel order, its phhp slots over the rows (a,i) in
    // mod-channel order
    std::vector<int32_t> blk(np, 0), blk2(np, 0), row_start_pp(npp, -1), ph_start(nslot, -1);
    for (const DChan &ch : c->chans)
        for (int AB = 0; AB < ch.pp; AB++) {
            const int a = ppl[ch.pp_off + AB].x - nh;
            row_start_pp[ch.pp_off + AB] = blk[a];
            blk[a] += ch.hh;
        }
    for (const DMod &m : c->mods)
        for (int q = 0; q < m.ph; q++) {
            const int a = phl[m.ph_off + q].x - nh, i = phl[m.ph_off + q].y;
            ph_start[slot(a + nh, i)] = blk2[a];
            blk2[a] += m.hp;
        }
    std::vector<int32_t> tr_base(np + 1, 0);
    int32_t max_blk = 0;
    for (int a = 0; a < np; a++) {
        if (blk[a] != blk2[a])
            return ccd::fail(CCD_E_INVALID, "particle %d has %d pphh and %d phhp elements", a + nh, blk[a], blk2[a]);
        max_blk = std::max(max_blk, blk[a]);
        tr_base[a + 1] = tr_base[a] + blk[a];
    }
    if (tr_base[np] != c->nt2) return ccd::fail(CCD_E_INVALID, "particle blocks do not cover T2");
    if (max_blk > 65535 || (size_t)(max_blk + 1) * sizeof(double) > 64 * 1024)
        return ccd::fail(CCD_E_INVALID, "a particle block of %d elements exceeds the transpose kernel's LDS", max_blk);
    c->tr_lds = (size_t)(max_blk + 1) * sizeof(double);   // plus the spare slot
    std::vector<int32_t> addr_pp(c->nt2, -1), addr_ph(c->nt2, -1);
    for (const DChan &ch : c->chans)
        for (int AB = 0; AB < ch.pp; AB++) {
            const int a = ppl[ch.pp_off + AB].x - nh;
            for (int IJ = 0; IJ < ch.hh; IJ++)
                addr_pp[tr_base[a] + row_start_pp[ch.pp_off + AB] + IJ] = (int32_t)(ch.off_t + (int64_t)AB * ch.hh + IJ);
        }
    for (const DMod &m : c->mods)
        for (int q = 0; q < m.ph; q++) {
            const int a = phl[m.ph_off + q].x - nh, i = phl[m.ph_off + q].y;
            for (int JB = 0; JB < m.hp; JB++)
                addr_ph[tr_base[a] + ph_start[slot(a + nh, i)] + JB] = (int32_t)(m.off_ph + (int64_t)q * m.hp + JB);
        }
    std::vector<uint16_t> perm_ph(c->nt2), perm_pp(c->nt2);
    std::vector<uint8_t> seen(c->nt2, 0);
    for (const DChan &ch : c->chans)
        for (int AB = 0; AB < ch.pp; AB++) {
            const int a = ppl[ch.pp_off + AB].x, b = ppl[ch.pp_off + AB].y, pa = a - nh;
            for (int IJ = 0; IJ < ch.hh; IJ++) {
                const int i = hhl[ch.hh_off + IJ].x, j = hhl[ch.hh_off + IJ].y;
                const int64_t s1 = slot(a, i), s2 = slot(b, j);
                const int32_t m1 = s1 < 0 ? -1 : mod_of_ph[s1], m2 = s2 < 0 ? -2 : mod_of_hp[s2];
                if (m1 < 0 || m1 != m2)
                    return ccd::fail(CCD_E_INVALID, "T2 element (%d,%d,%d,%d) has no consistent mod channel", a, b,
                                     i, j);
                const int32_t sph = ph_start[s1] + colidx[s2];
                const int32_t spp = row_start_pp[ch.pp_off + AB] + IJ;
                if (sph < 0 || sph >= blk[pa] || spp < 0 || spp >= blk[pa] || seen[tr_base[pa] + sph])
                    return ccd::fail(CCD_E_INVALID, "inconsistent pphh/phhp correspondence at (%d,%d,%d,%d)", a, b,
                                     i, j);
                seen[tr_base[pa] + sph] = 1;
                perm_ph[tr_base[pa] + sph] = (uint16_t)spp;  // phhp slot <- pphh slot
                perm_pp[tr_base[pa] + spp] = (uint16_t)sph;  // pphh slot <- phhp slot
            }
        }
    for (int64_t k = 0; k < c->nt2; k++)
        if (addr_pp[k] < 0 || addr_ph[k] < 0) return ccd::fail(CCD_E_INVALID, "transposition plan has a hole");

    // ---- tiles: packed-row tiles; ladder tiles = row tiles x k segments, grouped
    // by their 16-column tile count (channel order kept inside a group)
    for (int k = 0; k < (int)c->chans.size(); k++)
        for (int r = 0; r < c->chans[k].npk; r += LROWS)
            c->rtiles.push_back(DTile{k, r, std::min(LROWS, c->chans[k].npk - r), 0});
    for (int nt = 1; nt <= NTMAX; nt++) {
        for (int k = 0; k < (int)c->chans.size(); k++) {
            const DChan &ch = c->chans[k];
            if ((ch.hh + 15) / 16 != nt) continue;
            for (int sgm = 0; sgm < ch.nseg; sgm++)
                for (int r = 0; r < ch.npk; r += LROWS)
                    c->ltiles.push_back(DTile{k, r, std::min(LROWS, ch.npk - r), sgm});
        }
        c->lt_off[nt] = (int)c->ltiles.size();
    }
    // Block runs of the ladder tiles: per class the plan's block count and
    // assignment (plan_ladder_tiles), uploaded with the other tables below.
    HIP_TRY(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    c->ncu = std::max(1, c->ncu);
    c->ltiles0 = c->ltiles;
    {
        int base = 0;
        for (int nt = 1; nt <= NTMAX; nt++) {
            c->lb_base[nt] = base;
            base += c->lt_off[nt] - c->lt_off[nt - 1] + 1;
        }
        c->lbptr.assign((size_t)base, 0);
    }
    default_plan(c);
    plan_ladder_tiles(c);
    for (int wide = 0; wide < 2; wide++) {   // channels with hh <= 80 first
        for (int k = 0; k < (int)c->chans.size(); k++)
            if ((c->chans[k].hh > 80) == (wide == 1))
                for (int r = 0; r < c->chans[k].npk; r += WROWS)
                    c->ttiles.push_back(DTile{k, r, std::min(WROWS, c->chans[k].npk - r), 0});
        if (wide == 0) c->n_tt_narrow = (int)c->ttiles.size();
    }
    for (int k = 0; k < (int)c->chans.size(); k++) {
        const DChan &ch = c->chans[k];
        for (int r = 0; r < ch.hh; r += 16)
            for (int q = 0; q < ch.hh; q += 16) {
                for (int sg = 0; sg < ch.nxs; sg++) c->xtiles.push_back(DTile{k, r, sg, q});
                if (ch.nxs > 1) c->xftiles.push_back(DTile{k, r, 16, q});
            }
        const int npairs = ch.npk + ch.ndiag, per = std::max(1, kThreads / ch.hh);
        for (int r = 0; r < npairs; r += per) {
            c->utiles.push_back(DTile{k, r, std::min(per, npairs - r), (int32_t)c->n_cp});
            c->n_cp += ch.hh;
        }
    }
    if (c->n_cp >= (int64_t)INT32_MAX) return ccd::fail(CCD_E_INVALID, "too many column partials");
    c->ulo.assign(nprob + 1, 0);   // update tiles follow the channels, so each solve's are contiguous
    for (const DTile &t : c->utiles) c->ulo[c->chans[t.chan].prob + 1]++;
    for (int q = 0; q < nprob; q++) c->ulo[q + 1] += c->ulo[q];

    // ---- reduction lists: chi_i over column partials, psi_a over row sums
    std::vector<std::vector<int32_t>> lists((size_t)nsp);
    for (const DTile &t : c->utiles) {
        const DChan &ch = c->chans[t.chan];
        for (int IJ = 0; IJ < ch.hh; IJ++) lists[hhl[ch.hh_off + IJ].x].push_back(t.aux + IJ);
    }
    for (const DChan &ch : c->chans)
        for (int AB = 0; AB < ch.pp; AB++) lists[ppl[ch.pp_off + AB].x].push_back(ch.pp_off + AB);
    std::vector<int32_t> red_ptr(1, 0), red_idx;
    for (auto &v : lists) {
        red_idx.insert(red_idx.end(), v.begin(), v.end());
        red_ptr.push_back((int32_t)red_idx.size());
    }
    // the update stores the column partials at their slots and the row sums
    // after them (n_cp + row; rows of channels without hole pairs have no
    // sum and are in no list), so the lists hold places in that array
    if (c->n_cp + npp >= (int64_t)INT32_MAX) return ccd::fail(CCD_E_INVALID, "too many partial sums");
    if ((int64_t)red_idx.size() > c->n_cp + npp || red_ptr[nh] != c->n_cp)
        return ccd::fail(CCD_E_INVALID, "reduction lists do not match the partial sums");
    for (size_t q = (size_t)red_ptr[nh]; q < red_idx.size(); q++) red_idx[q] += (int32_t)c->n_cp;
    c->n_red = (int64_t)red_idx.size();

    // ---- upload tables
    int rc;
    if ((rc = upload(c, &c->d_chans, c->chans)) || (rc = upload(c, &c->d_mods, c->mods)) ||
        (rc = upload(c, &c->d_rtiles, c->rtiles)) || (rc = upload(c, &c->d_ltiles, c->ltiles)) ||
        (rc = upload(c, &c->d_utiles, c->utiles)) ||
        (rc = upload(c, &c->d_xtiles, c->xtiles)) || (rc = upload(c, &c->d_xftiles, c->xftiles)) ||
        (rc = upload(c, &c->d_ttiles, c->ttiles)) || (rc = upload(c, &c->d_ritems, c->ritems)) ||
        (rc = upload(c, &c->d_rgroups, c->rgroups)) || (rc = upload(c, &c->d_lbptr, c->lbptr)) ||
        (rc = upload(c, &c->d_t2_in, t2_in)) || (rc = upload(c, &c->d_t2_out, t2_out)) ||
        (rc = upload(c, &c->d_pp_old, pp_old)) ||
        (rc = upload(c, &c->d_ppl, ppl)) ||
        (rc = upload(c, &c->d_hhl, hhl)) || (rc = upload(c, &c->d_phl, phl)) || (rc = upload(c, &c->d_hpl, hpl)) ||
        (rc = upload(c, &c->d_hh_old, hh_old)) || (rc = upload(c, &c->d_prow, prow)) ||
        (rc = upload(c, &c->d_prow_sw, prow_sw)) || (rc = upload(c, &c->d_red_ptr, red_ptr)) ||
        (rc = upload(c, &c->d_red_idx, red_idx)) || (rc = upload(c, &c->d_tr_base, tr_base)) ||
        (rc = upload(c, &c->d_addr_pp, addr_pp)) || (rc = upload(c, &c->d_addr_ph, addr_ph)) ||
        (rc = upload(c, &c->d_perm_ph, perm_ph)) || (rc = upload(c, &c->d_perm_pp, perm_pp)))
        return rc;
    std::vector<double> eps(p->eps, p->eps + nsp);
    if ((rc = upload(c, &c->d_eps, eps))) return rc;
    const size_t nt = (size_t)c->nt2;
    if ((rc = dalloc(c, &c->d_onebody, (size_t)nsp)) || (rc = dalloc(c, &c->d_t[0], nt)) ||
        (rc = dalloc(c, &c->d_t[1], nt)) || (rc = dalloc(c, &c->d_tph, nt)) || (rc = dalloc(c, &c->d_r, nt)) ||
        (rc = dalloc(c, &c->d_g, nt)) || (rc = dalloc(c, &c->d_vt, nt)) || (rc = dalloc(c, &c->d_va, nt)) ||
        (rc = dalloc(c, &c->d_dp, (size_t)c->n_dp)) || (rc = dalloc(c, &c->d_tas, (size_t)c->n_tp)) ||
        (rc = dalloc(c, &c->d_dsym, c->precision == CCD_PREC_FP32 ? (size_t)c->n_sym : (size_t)1)) ||
        (rc = dalloc(c, &c->d_vtp, (size_t)c->n_tp)) || (rc = dalloc(c, &c->d_x, (size_t)c->n_hhhh)) ||
        (rc = dalloc(c, &c->d_xp, (size_t)std::max<int64_t>(c->n_xp, 1))) ||

        (rc = dalloc(c, &c->d_vh, (size_t)c->n_hhhh)) ||
        (rc = dalloc(c, &c->d_vp, c->matrix_free ? (size_t)1 : (size_t)c->n_vp)) ||
        (rc = dalloc(c, &c->d_vb, (size_t)c->n_b)) || (rc = dalloc(c, &c->d_red, (size_t)std::max<int64_t>(c->n_cp + npp, 1))) ||
        (rc = dalloc(c, &c->d_partials, std::max(c->utiles.size(), (size_t)1))) ||
        (rc = dalloc(c, &c->d_scalar, (size_t)std::max(4, nprob))) ||
        (rc = dalloc(c, &c->d_zero, (size_t)kZeroDoubles)) || (rc = dalloc(c, &c->d_ctl, 1)) || (rc = dalloc(c, &c->d_ticket, 1)) ||
        (rc = dalloc(c, &c->d_elog, (size_t)kElogCap * nprob)) || (rc = upload(c, &c->d_hlo, bi.hlo)) ||
        (rc = upload(c, &c->d_pprob, std::vector<int32_t>(sprob.begin() + nh, sprob.end()))))
        return rc;
    // every sum the update reads is written by the ladder of the same
    // iteration; zero for the solves of a batch that stop before their first
    HIP_TRY(hipMemset(c->d_dsym, 0, sizeof(float) * (c->precision == CCD_PREC_FP32 ? (size_t)std::max<int64_t>(c->n_sym, 1) : 1)));
    HIP_TRY(hipHostMalloc((void **)&c->h_ctl, sizeof(LoopCtl), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&c->h_ctl2, sizeof(LoopCtl), hipHostMallocDefault));
    *c->h_ctl = LoopCtl{};
    *c->h_ctl2 = LoopCtl{};
    if (nprob > 1) {
        if ((rc = dalloc(c, &c->d_pctl, (size_t)nprob))) return rc;
        c->lgen.pctl = c->d_pctl;   // the ladder skips stopped solves
        HIP_TRY(hipHostMalloc((void **)&c->h_pctl, sizeof(LoopCtl) * nprob, hipHostMallocDefault));
    } else {
        c->d_pctl = c->d_ctl;
    }
    HIP_TRY(hipMemset(c->d_ticket, 0, sizeof(int)));
    HIP_TRY(hipMemset(c->d_zero, 0, sizeof(double) * kZeroDoubles));
    HIP_TRY(hipHostMalloc((void **)&c->h_scalar, 4 * sizeof(double), hipHostMallocDefault));

    // ---- interaction
    if (p->gen_kind != CCD_GEN_NONE) {
        std::vector<int32_t> qn(p->qn, p->qn + 5 * (size_t)nsp);
        if ((rc = upload(c, &c->d_qn, qn))) return rc;
        Gen G{p->gen_kind, p->gen_nmax, p->gen_g, c->d_qn, nullptr, nullptr};
        if (p->gen_kind == kGenBatch) {   // per-solve parameters; merge_problems checked the states
            std::vector<double> tab(p->gen_table, p->gen_table + bi.table_len);
            if ((rc = upload(c, &c->d_table, tab)) || (rc = upload(c, &c->d_gpart, bi.parts))) return rc;
            G.table = c->d_table;
            G.part = c->d_gpart;
        } else if (p->gen_kind == CCD_GEN_ELECTRON_GAS) {
            for (int s = 0; s < nsp; s++)
                for (int k = 0; k < 3; k++)
                    if (std::abs(p->qn[5 * s + k]) > p->gen_nmax)
                        return ccd::fail(CCD_E_INVALID, "state %d has |n| > gen_nmax", s);
            const size_t w = 4 * (size_t)p->gen_nmax + 1;
            std::vector<double> tab(p->gen_table, p->gen_table + w * w * w);
            if ((rc = upload(c, &c->d_table, tab))) return rc;
            G.table = c->d_table;
        } else if (p->gen_kind == CCD_GEN_INF_MATTER) {
            std::vector<double> tab(p->gen_table, p->gen_table + 3);
            if ((rc = upload(c, &c->d_table, tab))) return rc;
            G.table = c->d_table;
        }
        if (c->matrix_free) {
            if ((rc = build_generated_ladder(c, p, bi, ppl, prow))) return rc;
            c->lgen.G = G;
        } else if (!c->rtiles.empty()) {
            hipLaunchKernelGGL(k_gen_vp, dim3((unsigned)c->rtiles.size()), dim3(kThreads), 0, c->stream,
                               c->d_rtiles, c->d_chans, c->d_ppl, c->d_prow, G, c->d_vp);
            if (c->precision == CCD_PREC_FP32) {   // the element the generated FP32 ladders build
                if ((rc = dalloc(c, &c->d_vp32, (size_t)std::max<int64_t>(c->n_vp, 1)))) return rc;
                hipLaunchKernelGGL(k_gen_vp32, dim3((unsigned)c->rtiles.size()), dim3(kThreads), 0, c->stream,
                                   c->d_rtiles, c->d_chans, c->d_ppl, c->d_prow, G, c->d_vp32);
            }
        }
        if (!c->chans.empty())
            hipLaunchKernelGGL(k_gen_direct_small, dim3((unsigned)c->chans.size()), dim3(kThreads), 0,
                               c->stream, c->d_chans, c->d_ppl, c->d_hhl, G, c->d_vt, c->d_vh);
        if (!c->mods.empty())
            hipLaunchKernelGGL(k_gen_mod, dim3((unsigned)c->mods.size()), dim3(kThreads), 0, c->stream,
                               c->d_mods, c->d_phl, c->d_hpl, G, c->d_va, c->d_vb);
        if ((rc = check_launch("generate V"))) return rc;
    } else {
        std::vector<int64_t> o_hhhh(p->n_chan + 1, 0), o_pppp(p->n_chan + 1, 0), o_b(p->n_mod + 1, 0);
        for (int k = 0; k < p->n_chan; k++) {
            o_hhhh[k + 1] = o_hhhh[k] + (int64_t)p->hh_dim[k] * p->hh_dim[k];
            o_pppp[k + 1] = o_pppp[k] + (int64_t)p->pp_dim[k] * p->pp_dim[k];
        }
        for (int k = 0; k < p->n_mod; k++) o_b[k + 1] = o_b[k] + (int64_t)p->hp_dim[k] * p->hp_dim[k];
        int ci = 0;
        std::vector<double> blkv;
        for (int k = 0; k < p->n_chan; k++) {
            if (!(p->pp_dim[k] > 0 && p->hh_dim[k] > 0)) continue;
            const DChan &ch = c->chans[ci++];
            // The packed ladder needs V^pppp antisymmetric in both pairs (exactly,
            // as calcTBME produces it); check, then keep the a<b, c<d block.
            // (V is in the caller's row order; r, q, rs and prow are internal rows)
            const double *V = p->v_pppp + o_pppp[k];
            const int P = ch.pp;
            const int32_t *old = pp_old.data() + ch.pp_off;
            for (int r = 0; r < P; r++) {
                const int rs = pp_swap[ch.pp_off + r];
                for (int q = 0; q < P; q++) {
                    const double v = V[(size_t)old[r] * P + old[q]];
                    if (V[(size_t)old[rs] * P + old[q]] != -v || V[(size_t)old[r] * P + old[pp_swap[ch.pp_off + q]]] != -v)
                        return ccd::fail(CCD_E_INVALID, "V^pppp of channel %d is not antisymmetric", k);
                }
            }
            blkv.assign((size_t)ch.npk * ch.ldv, 0.0);
            for (int u = 0; u < ch.npk; u++)
                for (int w = 0; w < ch.npk; w++)
                    blkv[(size_t)u * ch.ldv + w] = V[(size_t)old[prow[ch.pr_off + u]] * P + old[prow[ch.pr_off + w]]];
            HIP_TRY(hipMemcpy(c->d_vp + ch.off_vp, blkv.data(), sizeof(double) * blkv.size(), hipMemcpyHostToDevice));
            // V^hhhh in the internal column order on both sides
            const int H = ch.hh;
            const int32_t *hold = hh_old.data() + ch.hh_off;
            std::vector<double> vh((size_t)H * H);
            for (int r = 0; r < H; r++)
                for (int q = 0; q < H; q++) vh[(size_t)r * H + q] = p->v_hhhh[o_hhhh[k] + (int64_t)hold[r] * H + hold[q]];
            HIP_TRY(hipMemcpy(c->d_vh + ch.off_hhhh, vh.data(), sizeof(double) * vh.size(), hipMemcpyHostToDevice));
        }
        // V^hhpp has the pphh element count and offsets; stage through d_g, then transpose
        HIP_TRY(hipMemcpy(c->d_g, p->v_hhpp, sizeof(double) * nt, hipMemcpyHostToDevice));
        if (!c->chans.empty())
            hipLaunchKernelGGL(k_transpose_hhpp, dim3((unsigned)c->chans.size()), dim3(kThreads), 0, c->stream,
                               c->d_chans, c->d_pp_old, c->d_hh_old, c->d_g, c->d_vt);
        if ((rc = check_launch("k_transpose_hhpp"))) return rc;
        HIP_TRY(hipMemcpy(c->d_va, p->v_modA, sizeof(double) * nt, hipMemcpyHostToDevice));
        int mi = 0;
        for (int k = 0; k < p->n_mod; k++) {
            if (!(p->ph_dim[k] > 0 && p->hp_dim[k] > 0)) continue;
            const DMod &m = c->mods[mi++];
            HIP_TRY(hipMemcpy(c->d_vb + m.off_b, p->v_modB + o_b[k], sizeof(double) * m.hp * (size_t)m.hp,
                              hipMemcpyHostToDevice));
        }
    }
    if (!c->rtiles.empty()) {
        hipLaunchKernelGGL(k_pack_vt, dim3((unsigned)c->rtiles.size()), dim3(kThreads), 0, c->stream, c->d_rtiles,
                           c->d_chans, c->d_prow, c->d_vt, c->d_vtp);
        if ((rc = check_launch("k_pack_vt"))) return rc;
    }
    if (c->precision == CCD_PREC_FP32) {   // FP32 copies for the mixed-precision iterations
        auto to_f32 = [&](const double *src, float **dst, int64_t n) {
            int r = dalloc(c, dst, (size_t)std::max<int64_t>(n, 1));
            if (r || n == 0) return r;
            const unsigned g = (unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, 4096);
            hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(kThreads), 0, c->stream, src, *dst, n);
            return check_launch("k_to_f32");
        };
        // V^pppp: generated contexts built theirs above (or generate it in the ladder)
        if ((rc = to_f32(c->d_va, &c->d_va32, c->nt2)) || (rc = to_f32(c->d_vb, &c->d_vb32, c->n_b)) ||
            (rc = to_f32(c->d_vt, &c->d_vt32, c->nt2)) ||
            (!c->d_vp32 && (rc = to_f32(c->d_vp, &c->d_vp32, c->matrix_free ? 0 : c->n_vp))))
            return rc;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

// Arm the device loop state.  fresh: iteration count 0 and E_prev = 100.0
// (src/main.cpp:494); otherwise the count and E_prev carry on.  tol < 0 and
// max_iters = INT_MAX give a free run that never stops by itself.
// A batch arms every solve this way; a solve that stopped other than by
// converging stays stopped when the loop carries on.  The batch's own state
// counts iterations and stops once every solve has.
int arm_loop(ccd_ctx *c, double tol, int max_iters, bool fresh) {
    HIP_TRY(hipStreamSynchronize(c->stream));   // the staging buffers may still be in use
    const LoopCtl start{0, 0, CCD_STOP_CONVERGED, max_iters, 100.0, tol, kElogCap, 0};
    if (c->nprob == 1) {
        if (fresh) {
            *c->h_ctl = start;
        } else {
            HIP_TRY(hipMemcpy(c->h_ctl, c->d_ctl, sizeof(LoopCtl), hipMemcpyDeviceToHost));
            c->h_ctl->done = 0;
            c->h_ctl->reason = CCD_STOP_CONVERGED;
            c->h_ctl->tol = tol;
            c->h_ctl->max_iters = max_iters;
        }
        HIP_TRY(hipMemcpyAsync(c->d_ctl, c->h_ctl, sizeof(LoopCtl), hipMemcpyHostToDevice, c->stream));
        return CCD_OK;
    }
    const size_t pb = sizeof(LoopCtl) * (size_t)c->nprob;
    if (fresh) {
        for (int q = 0; q < c->nprob; q++) c->h_pctl[q] = start;
        *c->h_ctl = LoopCtl{0, 0, CCD_STOP_CONVERGED, INT_MAX, 100.0, -1.0, kElogCap, 0};
    } else {
        HIP_TRY(hipMemcpy(c->h_pctl, c->d_pctl, pb, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(c->h_ctl, c->d_ctl, sizeof(LoopCtl), hipMemcpyDeviceToHost));
        bool running = false;
        for (int q = 0; q < c->nprob; q++) {
            LoopCtl &l = c->h_pctl[q];
            if (l.reason == CCD_STOP_CONVERGED) {
                l.done = 0;
                l.tol = tol;
                l.max_iters = max_iters;
            }
            running |= !l.done;
        }
        c->h_ctl->done = running ? 0 : 1;
    }
    HIP_TRY(hipMemcpyAsync(c->d_pctl, c->h_pctl, pb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_ctl, c->h_ctl, sizeof(LoopCtl), hipMemcpyHostToDevice, c->stream));
    return CCD_OK;
}

// Per-solve loop state of a batch into h_pctl (after a run_loop).
int read_pctl(ccd_ctx *c) {
    if (c->nprob == 1) return CCD_OK;
    HIP_TRY(hipMemcpyAsync(c->h_pctl, c->d_pctl, sizeof(LoopCtl) * (size_t)c->nprob, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int arm_free_run(ccd_ctx *c) {
    int rc = arm_loop(c, -1.0, INT_MAX, true);
    return rc ? rc : sync_state(c);   // T^ph in the type the next iterations read
}

// A device loop stopped: the T remap and the psi blocks of the reduction of
// its last iteration ran beside the reduction block that set the stop flag,
// and those that started after it returned at entry.  The amplitudes and
// their antisymmetrised rows (written by the update before) are whole;
// T^ph, chi and psi are rebuilt from them before the next iteration.
void mark_stopped(ccd_ctx *c) {
    c->tph_prec = kStale;
    c->onebody_stale = true;
}

// The loop state of a free run, with no rebuild (the end of a solve: the
// next entry point that iterates calls arm_free_run).
int end_solve(ccd_ctx *c) { return arm_loop(c, -1.0, INT_MAX, true); }

// Queue iterations in batches without waiting for each energy, one batch
// ahead: the loop state after batch k is read while batch k+1 runs.
// Iterations queued past the stopping point do nothing.  Returns with the
// loop state in h_ctl and the current amplitudes set to those of the last
// real iteration.
int run_loop(ccd_ctx *c, double mixing, int batch) {
    const int cur0 = c->cur, count0 = c->h_ctl->count;
    LoopCtl *slot[2] = {c->h_ctl, c->h_ctl2};
    const hipStream_t cs = ctl_stream(c);
    int rc;
    auto queue = [&](int k) -> int {
        if (int rq = run_batch(c, mixing, batch)) return rq;
        HIP_TRY(hipMemcpyAsync(slot[k], c->d_ctl, sizeof(LoopCtl), hipMemcpyDeviceToHost, cs));
        HIP_TRY(hipEventRecord(c->ctl_ev[k], cs));
        return CCD_OK;
    };
    if ((rc = begin_run(c)) || (rc = queue(0))) return rc;
    int k = 0;
    for (;; k ^= 1) {
        if ((rc = queue(k ^ 1))) return rc;
        HIP_TRY(hipEventSynchronize(c->ctl_ev[k]));
        if (slot[k]->done) break;
    }
    // the batch queued after the stop has run (returning at entry) and left
    // the same state in the other slot
    if ((rc = end_run(c))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (slot[k] != c->h_ctl) *c->h_ctl = *slot[k];
    c->cur = (cur0 + (c->h_ctl->count - count0)) & 1;
    mark_stopped(c);
    return CCD_OK;
}

// Energies of solve q (0 for a single solve): the first min(cap, count)
// entries of its log.
int copy_log(ccd_ctx *c, double *energies, int cap, int count, int q = 0) {
    const int n = std::min(std::min(cap, count), kElogCap);
    if (energies && n > 0) {
        HIP_TRY(hipMemcpyAsync(energies, c->d_elog + (size_t)q * kElogCap, sizeof(double) * n,
                               hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return CCD_OK;
}

int zero_state(ccd_ctx *c) {
    const size_t nt = sizeof(double) * (size_t)c->nt2;
    HIP_TRY(hipMemsetAsync(c->d_t[0], 0, nt, c->stream));
    HIP_TRY(hipMemsetAsync(c->d_t[1], 0, nt, c->stream));
    HIP_TRY(hipMemsetAsync(c->d_tph, 0, nt, c->stream));   // zero in either element type
    c->tph_prec = c->iter_prec;
    HIP_TRY(hipMemsetAsync(c->d_tas, 0, sizeof(double) * (size_t)std::max<int64_t>(c->n_tp, 1), c->stream));
    c->tas_prec = c->iter_prec;
    HIP_TRY(hipMemsetAsync(c->d_onebody, 0, sizeof(double) * c->nsp, c->stream));
    c->onebody_stale = false;
    c->ob_prec = kAnyPrec;
    c->cur = 0;
    return arm_free_run(c);
}

// Load host amplitudes as the current T and derive T^ph, the antisymmetrised
// packed rows, chi and psi from them.
int set_state(ccd_ctx *c, const double *t2) {
    if (int rc0 = arm_loop(c, -1.0, INT_MAX, true)) return rc0;   // a free run; everything derived is rebuilt here
    double *t = c->d_t[c->cur];
    int rc;
    if ((rc = t2_upload(c, t2, t, c->d_g))) return rc;
    if ((rc = launch_tremap(c, t, c->stream)) || (rc = launch_antisym(c, t)) ||
        (rc = launch_update<false>(c, t, nullptr, 0.0, c->stream)) || (rc = launch_reduce(c, c->stream)))
        return rc;
    c->onebody_stale = false;
    return CCD_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

}  // extern "C"

namespace {

int create_ctx(const ccd_problem *prob, const BatchInfo &bi, int precision, ccd_ctx **out) {
    int rc;
    const bool matrix_free = (precision & CCD_MATRIX_FREE) != 0;
    precision &= ~CCD_MATRIX_FREE;
    if (precision != CCD_PREC_FP64 && precision != CCD_PREC_FP32)
        return ccd::fail(CCD_E_INVALID, "precision %d is not available", precision);
    if (matrix_free && prob->gen_kind == CCD_GEN_NONE)
        return ccd::fail(CCD_E_INVALID, "the matrix-free mode needs an on-device generator (gen_kind)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ccd::fail(CCD_E_HIP, "no HIP device available");
    ccd_ctx *c = new (std::nothrow) ccd_ctx();
    if (!c) return ccd::fail(CCD_E_NOMEM, "out of host memory");
    c->precision = precision;
    c->iter_prec = precision;
    c->matrix_free = matrix_free;
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->aux[0], hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->aux[1], hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->aux[2], hipStreamNonBlocking) != hipSuccess) {
        free_ctx(c);
        return ccd::fail(CCD_E_HIP, "could not create the HIP streams");
    }
    for (hipEvent_t &e : c->sync_ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            free_ctx(c);
            return ccd::fail(CCD_E_HIP, "could not create HIP events");
        }
    for (hipEvent_t &e : c->ctl_ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            free_ctx(c);
            return ccd::fail(CCD_E_HIP, "could not create HIP events");
        }
    if ((rc = init_scratch(c->device))) {
        free_ctx(c);
        return rc;
    }
    if ((rc = allow_ring_lds_all()) || (rc = allow_ladder_lds_all())) {
        free_ctx(c);
        return rc;
    }
    try {
        rc = build(c, prob, bi);
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

// The schedule tuner: a free run of a few batches under each schedule (the
// first batch untimed: it captures the graph), HIP events on the main
// stream around the timed batches; the fastest is kept, and the amplitudes
// go back to T2 = 0.  Every schedule gives the same results bit for bit
// (the same kernels in a dependency-respecting order, every reduction in a
// fixed order), so the choice is about time only.
int tune_schedule(ccd_ctx *c) {
    static const int cand[3] = {CCD_SCHED_STREAMS, CCD_SCHED_SERIAL, CCD_SCHED_SERIAL_GRAPH};
    constexpr int kTimed = 2;   // batches
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return ccd::fail(CCD_E_HIP, "could not create HIP events");
    }
    int rc = CCD_OK, pick = CCD_SCHED_STREAMS;
    double best = 0.0;
    for (int q = 0; q < 3 && !rc; q++) {
        c->sched = cand[q];
        rc = arm_free_run(c);
        if (!rc) rc = run_n(c, 0.5, kSolveBatch);
        hipError_t e = rc ? hipSuccess : hipEventRecord(a, c->stream);
        if (!rc && e == hipSuccess) rc = run_n(c, 0.5, kTimed * kSolveBatch);
        if (!rc && e == hipSuccess) e = hipEventRecord(b, c->stream);
        if (!rc && e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.f;
        if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (!rc && e != hipSuccess) rc = ccd::fail(CCD_E_HIP, "schedule timing failed: %s", hipGetErrorString(e));
        if (rc) break;
        c->sched_us[q] = 1e3 * ms / (kTimed * kSolveBatch);
        if (q == 0 || c->sched_us[q] < best) {
            best = c->sched_us[q];
            pick = cand[q];
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    c->sched = pick;
    if (!rc && pick != CCD_SCHED_SERIAL_GRAPH) rc = drop_graphs(c);
    if (!rc) rc = zero_state(c);
    return rc;
}

// Device time of one launch of ladder class nt (microseconds), averaged over
// `reps` launches back to back on the main stream after one untimed launch.
int time_class(ccd_ctx *c, int nt, int reps, double *us) {
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return ccd::fail(CCD_E_HIP, "could not create HIP events");
    }
    int rc = launch_ladder_v(c, nt, nt, c->stream);
    hipError_t e = rc ? hipSuccess : hipEventRecord(a, c->stream);
    for (int r = 0; r < reps && !rc && e == hipSuccess; r++) rc = launch_ladder_v(c, nt, nt, c->stream);
    if (!rc && e == hipSuccess) e = hipEventRecord(b, c->stream);
    if (!rc && e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (!rc && e != hipSuccess) rc = ccd::fail(CCD_E_HIP, "ladder timing failed: %s", hipGetErrorString(e));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!rc) *us = 1e3 * ms / reps;
    return rc;
}

// Microseconds per iteration of the context's schedule (a free run: one
// untimed batch, then kTimed timed ones).
int time_iterations(ccd_ctx *c, double *us) {
    constexpr int kTimed = 4;
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return ccd::fail(CCD_E_HIP, "could not create HIP events");
    }
    int rc = arm_free_run(c);
    if (!rc) rc = run_n(c, 0.5, kSolveBatch);
    hipError_t e = rc ? hipSuccess : hipEventRecord(a, c->stream);
    if (!rc && e == hipSuccess) rc = run_n(c, 0.5, kTimed * kSolveBatch);
    if (!rc && e == hipSuccess) e = hipEventRecord(b, c->stream);
    if (!rc && e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (!rc && e != hipSuccess) rc = ccd::fail(CCD_E_HIP, "iteration timing failed: %s", hipGetErrorString(e));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!rc) *us = 1e3 * ms / (kTimed * kSolveBatch);
    return rc;
}

int plan_valid(const ccd_ladder_plan &p) {
    for (int n = 0; n < NTMAX; n++) {
        for (int q = 0; q < 2; q++)
            if (p.kind[q][n] < CCD_LK_DEFAULT || p.kind[q][n] > CCD_LK_BF16X3W)
                return ccd::fail(CCD_E_INVALID, "ladder plan: unknown kernel kind %d (class %d)", p.kind[q][n], n + 1);
        if (p.kind[CCD_PREC_FP64][n] >= CCD_LK_BF16X3)
            return ccd::fail(CCD_E_INVALID, "ladder plan: BF16x3 products are for FP32 iterations (class %d)", n + 1);
        if (n > 0 && p.kind[CCD_PREC_FP32][n] >= CCD_LK_BF16X3SYM)
            return ccd::fail(CCD_E_INVALID, "ladder plan: kernel kind %d is for class 1 (class %d)",
                             p.kind[CCD_PREC_FP32][n], n + 1);
        if (p.blocks_per_cu[n] < 0 || p.blocks_per_cu[n] > 16)
            return ccd::fail(CCD_E_INVALID, "ladder plan: %d blocks per compute unit (class %d)", p.blocks_per_cu[n],
                             n + 1);
        if (p.assign[n] < CCD_LA_MODEL || p.assign[n] > CCD_LA_XCD)
            return ccd::fail(CCD_E_INVALID, "ladder plan: unknown assignment %d (class %d)", p.assign[n], n + 1);
    }
    if (p.partition != CCD_LP_SPLIT && p.partition != CCD_LP_ONE)
        return ccd::fail(CCD_E_INVALID, "ladder plan: unknown partition %d", p.partition);
    if (p.ring_fp32 != CCD_RK_FP32 && p.ring_fp32 != CCD_RK_BF16X3)
        return ccd::fail(CCD_E_INVALID, "ladder plan: unknown ring kind %d", p.ring_fp32);
    return CCD_OK;
}

// The ladder tuner (ccd_tune_ladder).  Per class with tiles, a coordinate
// search -- kernel kind, then blocks per compute unit, then assignment --
// each candidate timed over repeated launches of the class alone.  Classes
// timed alone do not show how they share the GPU with each other and with
// the ring and the remaps in the four-stream schedule, so the final choice
// is made on whole iterations in the context's schedule, among: the plan
// the tuner started from, the per-class result, the starting plan and the
// per-class one with each other allowed kind in every class, each under
// both partitions.  The
// amplitudes of two free iterations give the kernels realistic operands;
// T2 goes back to zero at the end.
int tune_ladder(ccd_ctx *c, int flags) {
    constexpr int kReps = 5;
    const int prec = c->iter_prec;
    const ccd_ladder_plan start = c->plan;
    ccd_ladder_plan &p = c->plan;
    std::vector<int> kinds{CCD_LK_DEFAULT};
    if (prec == CCD_PREC_FP32 && (flags & CCD_TUNE_KINDS)) kinds.push_back(CCD_LK_MFMA16);
    if (prec == CCD_PREC_FP32 && (flags & CCD_TUNE_BF16X3)) kinds.push_back(CCD_LK_BF16X3);
    int rc = drop_graphs(c);
    if (!rc) rc = arm_free_run(c);
    if (!rc) rc = run_n(c, 0.5, 2);
    if (!rc) HIP_TRY(hipStreamSynchronize(c->stream));
    double us_class[NTMAX] = {};
    for (int nt = 1; nt <= NTMAX && !rc; nt++) {
        if (c->lt_off[nt] == c->lt_off[nt - 1]) continue;
        double best = 0.0;
        auto consider = [&](int32_t &field, const std::vector<int> &vals) -> int {
            int pick = field;
            for (size_t q = 0; q < vals.size(); q++) {
                field = vals[q];
                plan_class_tiles(c, nt);
                double us = 0.0;
                int r = upload_ladder_tiles(c);
                if (!r) r = time_class(c, nt, kReps, &us);
                if (r) return r;
                if (best == 0.0 || us < best) {
                    best = us;
                    pick = vals[q];
                }
            }
            field = pick;
            plan_class_tiles(c, nt);
            return upload_ladder_tiles(c);
        };
        std::vector<int> ck{CCD_LK_DEFAULT};   // class 1's 16x16x4 kernel differs from its default
        for (int k : kinds) {
            if (k == CCD_LK_BF16X3 || (k == CCD_LK_MFMA16 && nt == 1)) ck.push_back(k);
            if (k == CCD_LK_BF16X3 && nt == 1) {
                ck.push_back(CCD_LK_BF16X3SYM);
                ck.push_back(CCD_LK_BF16X3W);
            }
        }
        std::vector<int> bpcs{1, 2, 3, 4, 6};
        if (!c->matrix_free) bpcs.insert(bpcs.begin(), 0);
        if ((rc = consider(p.kind[prec][nt - 1], ck)) || (rc = consider(p.blocks_per_cu[nt - 1], bpcs)) ||
            (rc = consider(p.assign[nt - 1], {CCD_LA_CHANNEL, CCD_LA_LONGEST, CCD_LA_RUNS, CCD_LA_XCD})))
            break;
        us_class[nt - 1] = best;
    }
    // whole plans, whole iterations: each kind in every class on the
    // starting plan's blocks and assignments and on the per-class ones
    const ccd_ladder_plan per_class = p;
    std::vector<ccd_ladder_plan> cands{start, p};
    for (const ccd_ladder_plan *base : {&start, &per_class})
        for (int k : kinds) {
            if (k == CCD_LK_DEFAULT) continue;
            ccd_ladder_plan q = *base;
            for (int n = 0; n < NTMAX; n++) q.kind[prec][n] = k;
            cands.push_back(q);
            if (k == CCD_LK_BF16X3) {   // and class 1 on the wave-k kernels
                q.kind[prec][0] = CCD_LK_BF16X3SYM;
                cands.push_back(q);
                q.kind[prec][0] = CCD_LK_BF16X3W;
                cands.push_back(q);
            }
        }
    if (prec == CCD_PREC_FP32 && (flags & CCD_TUNE_RING_BF16X3)) {   // each of them with the BF16x3 ring
        const size_t nc = cands.size();
        for (size_t q = 0; q < nc; q++) {
            ccd_ladder_plan r = cands[q];
            r.ring_fp32 = CCD_RK_BF16X3;
            cands.push_back(r);
        }
    }
    double best = 0.0;
    ccd_ladder_plan pick = start;
    for (size_t q = 0; q < cands.size() && !rc; q++)
        for (int part = CCD_LP_SPLIT; part <= CCD_LP_ONE && !rc; part++) {
            p = cands[q];
            p.partition = part;
            plan_ladder_tiles(c);
            double us = 0.0;
            if ((rc = upload_ladder_tiles(c)) || (rc = drop_graphs(c)) || (rc = time_iterations(c, &us))) break;
            cands[q].us_partition[part] = us;
            if (best == 0.0 || us < best) best = us;
        }
    // whole iterations of plans this close differ by the noise of the
    // timing: the first candidate within 1% of the fastest (the per-class
    // result comes right after the starting plan)
    for (size_t q = 0; q < cands.size() && !rc; q++) {
        int part = CCD_LP_SPLIT;
        while (part <= CCD_LP_ONE && !(cands[q].us_partition[part] <= 1.01 * best)) part++;
        if (part <= CCD_LP_ONE) {
            pick = cands[q];
            pick.partition = part;
            break;
        }
    }
    if (!rc) {
        p = pick;
        for (int n = 0; n < NTMAX; n++) p.us[prec][n] = us_class[n];
        for (size_t q = 0; q < cands.size(); q++)   // the partitions' timings of the plan chosen
            if (std::memcmp(cands[q].kind, pick.kind, sizeof pick.kind) == 0 &&
                std::memcmp(cands[q].blocks_per_cu, pick.blocks_per_cu, sizeof pick.blocks_per_cu) == 0 &&
                std::memcmp(cands[q].assign, pick.assign, sizeof pick.assign) == 0 &&
                cands[q].ring_fp32 == pick.ring_fp32) {
                p.us_partition[0] = cands[q].us_partition[0];
                p.us_partition[1] = cands[q].us_partition[1];
            }
        plan_ladder_tiles(c);
        rc = upload_ladder_tiles(c);
    }
    if (!rc) rc = drop_graphs(c);
    if (!rc) rc = zero_state(c);
    return rc;
}

int batch_only(const ccd_ctx *c, const char *what) {
    return ccd::fail(CCD_E_STATE, "%s: context holds a batch of %d solves (use the ccd_batch_* calls)", what,
                     c->nprob);
}

}  // namespace

extern "C" {

int ccd_ctx_create(const ccd_problem *prob, int precision, ccd_ctx **out) {
    ccd::clear_error();
    if (!out) return ccd::fail(CCD_E_INVALID, "ccd_ctx_create: out is null");
    *out = nullptr;
    int rc = validate(prob);
    if (rc) return rc;
    return create_ctx(prob, single_batch(prob), precision, out);
}

int ccd_batch_create(const ccd_problem *const *probs, int n, int precision, ccd_ctx **out) {
    ccd::clear_error();
    if (!out) return ccd::fail(CCD_E_INVALID, "ccd_batch_create: out is null");
    *out = nullptr;
    if (!probs || n <= 0) return ccd::fail(CCD_E_INVALID, "ccd_batch_create: no problems");
    Merged m;
    int rc;
    try {
        rc = merge_problems(probs, n, m);
    } catch (const std::bad_alloc &) {
        rc = ccd::fail(CCD_E_NOMEM, "out of host memory while merging the batch");
    }
    if (rc) return rc;
    return create_ctx(&m.p, m.bi, precision, out);
}

int ccd_batch_info(const ccd_ctx *c, int *n_problems, int64_t *t2_offset) {
    ccd::clear_error();
    if (!c) return ccd::fail(CCD_E_INVALID, "null context");
    if (n_problems) *n_problems = c->nprob;
    if (t2_offset)
        for (int q = 0; q <= c->nprob; q++) t2_offset[q] = c->t2_lo[q];
    return CCD_OK;
}

int ccd_batch_mbpt2(ccd_ctx *c, double *corr) {
    ccd::clear_error();
    if (!c || !corr) return ccd::fail(CCD_E_INVALID, "ccd_batch_mbpt2: null argument");
    int rc;
    if (!c->utiles.empty()) {
        UpdArgs a = upd_args(c, nullptr, nullptr, 0.0);
        hipLaunchKernelGGL(k_mbpt2, dim3((unsigned)c->utiles.size()), dim3(kThreads), 0, c->stream, a,
                           c->d_partials);
        if ((rc = check_launch("k_mbpt2"))) return rc;
    }
    for (int q = 0; q < c->nprob; q++) {
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, c->stream, c->d_partials + c->ulo[q],
                           c->ulo[q + 1] - c->ulo[q], 0.25, c->d_scalar + q);
        if ((rc = check_launch("k_sum_partials"))) return rc;
    }
    HIP_TRY(hipMemcpyAsync(corr, c->d_scalar, sizeof(double) * c->nprob, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int ccd_batch_solve(ccd_ctx *c, double tol, int max_iters, double mixing, double *energies, int cap, int *n_iter,
                    int *stop) {
    ccd::clear_error();
    if (!c) return ccd::fail(CCD_E_INVALID, "null context");
    if (max_iters < 0 || cap < 0) return ccd::fail(CCD_E_INVALID, "max_iters or cap < 0");
    int rc = zero_state(c);
    if (rc || (rc = arm_loop(c, tol, max_iters, true)) || (rc = run_loop(c, mixing, kSolveBatch)) ||
        (rc = read_pctl(c)))
        return rc;
    for (int q = 0; q < c->nprob; q++) {
        const LoopCtl st = c->nprob == 1 ? *c->h_ctl : c->h_pctl[q];
        if ((rc = copy_log(c, energies ? energies + (size_t)q * cap : nullptr, cap, st.count, q))) return rc;
        if (n_iter) n_iter[q] = st.count;
        if (stop) stop[q] = st.reason;
    }
    return end_solve(c);
}

int ccd_batch_solve_mixed(ccd_ctx *c, double tol, double switch_tol, int max_iters, double mixing,
                          double *energies, int cap, int *n_iter, int *n_fp32, int *stop) {
    ccd::clear_error();
    if (!c) return ccd::fail(CCD_E_INVALID, "null context");
    if (c->precision != CCD_PREC_FP32)
        return ccd::fail(CCD_E_STATE, "ccd_batch_solve_mixed needs a context created with CCD_PREC_FP32");
    if (max_iters < 0 || cap < 0) return ccd::fail(CCD_E_INVALID, "max_iters or cap < 0");
    const int saved = c->iter_prec;
    std::vector<LoopCtl> phase1(c->nprob);
    int rc = zero_state(c);
    c->iter_prec = CCD_PREC_FP32;
    c->tph_prec = CCD_PREC_FP32;   // T^ph and tas are zero: the same in either type
    c->tas_prec = CCD_PREC_FP32;
    if (!rc) rc = arm_loop(c, std::max(tol, switch_tol), max_iters, true);
    if (!rc) rc = run_loop(c, mixing, kSolveBatch);
    if (!rc) rc = read_pctl(c);
    if (!rc)
        for (int q = 0; q < c->nprob; q++) phase1[q] = c->nprob == 1 ? *c->h_ctl : c->h_pctl[q];
    if (!rc && switch_tol > tol) {
        c->iter_prec = CCD_PREC_FP64;
        rc = arm_loop(c, tol, max_iters, false);
        if (!rc) rc = sync_state(c);
        if (!rc) rc = run_loop(c, mixing, kSolveBatch);
        if (!rc) rc = read_pctl(c);
    }
    c->iter_prec = saved;   // the derived state follows at the next iteration (sync_state)
    if (rc) return rc;
    for (int q = 0; q < c->nprob; q++) {
        const LoopCtl st = c->nprob == 1 ? *c->h_ctl : c->h_pctl[q];
        if ((rc = copy_log(c, energies ? energies + (size_t)q * cap : nullptr, cap, st.count, q))) return rc;
        if (n_iter) n_iter[q] = st.count;
        if (n_fp32) n_fp32[q] = phase1[q].count;
        if (stop) stop[q] = st.reason;
    }
    return end_solve(c);
}

int ccd_batch_iterate_n(ccd_ctx *c, int n, double mixing, double *energies) {
    ccd::clear_error();
    if (!c || n < 0) return ccd::fail(CCD_E_INVALID, "ccd_batch_iterate_n: bad argument");
    int rc = arm_free_run(c);
    if (rc || (rc = run_n(c, mixing, n))) return rc;
    HIP_TRY(hipMemcpyAsync(c->h_ctl, c->d_ctl, sizeof(LoopCtl), hipMemcpyDeviceToHost, c->stream));
    if ((rc = read_pctl(c))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->h_ctl->done) mark_stopped(c);   // every solve has stopped (NaN)
    if (c->nprob == 1 && c->h_ctl->count != n) {   // a NaN energy stops even a free run
        c->cur = (c->cur + c->h_ctl->count - n) & 1;
        return ccd::fail(CCD_E_STATE, "iteration %d produced a NaN energy", c->h_ctl->count);
    }
    for (int q = 0; q < c->nprob; q++) {
        // a solve of a batch whose energy turned NaN stopped there (its
        // amplitudes kept); the entries after its last iteration are NaN
        const int cnt = c->nprob == 1 ? n : c->h_pctl[q].count;
        double *eq = energies ? energies + (size_t)q * n : nullptr;
        if ((rc = copy_log(c, eq, n, cnt, q))) return rc;
        if (eq)
            for (int k = std::min(cnt, kElogCap); k < n; k++) eq[k] = std::numeric_limits<double>::quiet_NaN();
    }
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
        // one pass of the pppp ladder reads the packed V^pppp (unless
        // generated), the antisymmetrised amplitudes and X once and writes D_p
        // once (V^pppp, the amplitude rows and D_p in the arithmetic type of
        // the iteration, X in FP64)
        const double vb = ctx->iter_prec == CCD_PREC_FP64 ? 8.0 : 4.0;
        double b = 0.0;
        for (const DChan &ch : ctx->chans)
            b += (ctx->matrix_free ? 0.0 : vb * ch.npk * ch.npk) + 2.0 * vb * ch.npk * ch.hh +
                 8.0 * (double)ch.hh * ch.hh;
        *ladder_bytes = b;
    }
    return CCD_OK;
}

int ccd_kernel_work(const ccd_ctx *ctx, double *flops, double *bytes, double *elements, int n) {
    if (!ctx || n < 0) return ccd::fail(CCD_E_INVALID, "ccd_kernel_work: bad argument");
    double f[CCD_KT_COUNT] = {}, b[CCD_KT_COUNT] = {}, e[CCD_KT_COUNT] = {};
    const double cs = ctx->iter_prec == CCD_PREC_FP64 ? 8.0 : 4.0;   // contraction operands and outputs
    const double nt = (double)ctx->nt2;
    for (const DMod &m : ctx->mods) {
        const double ph = m.ph, hp = m.hp;
        f[CCD_KT_RING] += 2.0 * hp * hp * ph + 2.0 * ph * hp * hp;
        b[CCD_KT_RING] += cs * (hp * ph + hp * hp + 2.0 * ph * hp);   // V^A, V^B, T^ph in, R out
    }
    b[CCD_KT_REMAP_R] = 2.0 * cs * nt;
    b[CCD_KT_REMAP_T] = (8.0 + cs) * nt;   // FP64 amplitudes in, T^ph in the contractions' type out
    for (const DChan &ch : ctx->chans) {
        const double npk = ch.npk, hh = ch.hh, ppk = (double)ch.npk * ch.npk;
        f[CCD_KT_X] += 2.0 * hh * hh * npk;
        b[CCD_KT_X] += (8.0 + cs) * npk * hh + 16.0 * hh * hh;      // V^hhpp rows, T rows, V^hhhh in, X out
        f[CCD_KT_TX] += 2.0 * npk * hh * hh;
        b[CCD_KT_TX] += 2.0 * cs * npk * hh + 8.0 * hh * hh;        // T rows, X in, D part out
        const int nt16 = (ch.hh + 15) / 16;
        const int g = nt16 == 1 ? CCD_KT_LADDER1 : nt16 == 2 ? CCD_KT_LADDER2 : nt16 <= 4 ? CCD_KT_LADDER34
                                                                                          : CCD_KT_LADDER58;
        f[g] += 2.0 * ppk * hh;
        if (g == CCD_KT_LADDER1 && ch.sym >= 0 && ctx->iter_prec == CCD_PREC_FP32 &&
            ctx->plan.kind[CCD_PREC_FP32][0] == CCD_LK_BF16X3SYM) {
            // the symmetric kernel: tile J builds (or reads) the elements of its
            // rows from k = 64 J on, and writes the sums of the rows past it,
            // which the update reads
            double el = 0.0, tr = 0.0;
            for (int r0 = 0; r0 < ch.npk; r0 += LROWS) {
                el += (double)std::min(LROWS, ch.npk - r0) * (ch.npk - r0);
                tr += (double)std::max(0, ch.npk - r0 - LROWS) * hh;
            }
            b[g] += (ctx->matrix_free ? 0.0 : cs * el) + 2.0 * cs * npk * hh + 4.0 * tr;
            b[CCD_KT_UPDATE] += 4.0 * tr;
            if (ctx->matrix_free) e[g] += el;
        } else {
            b[g] += (ctx->matrix_free ? 0.0 : cs * ppk) + 2.0 * cs * npk * hh;   // V^pppp (stored), T rows in, D out
            if (ctx->matrix_free) e[g] += ppk;
        }
        // the update: D read once per packed row (one array in the algorithm)
        b[CCD_KT_UPDATE] += cs * npk * hh + cs * npk * hh;               // D in, T rows out
    }
    // per element: <ij||ab> and the ring term (both in the contraction type),
    // T_old in, T_new out
    b[CCD_KT_UPDATE] += (cs + 8.0 + cs + 8.0) * nt;
    b[CCD_KT_REDUCE] = 12.0 * (double)ctx->n_red + 8.0 * ctx->nsp;
    for (int k = 0; k < n && k < CCD_KT_COUNT; k++) {
        if (flops) flops[k] = f[k];
        if (bytes) bytes[k] = b[k];
        if (elements) elements[k] = e[k];
    }
    return CCD_OK;
}

int ccd_mbpt2(ccd_ctx *c, double *corr) {
    ccd::clear_error();
    if (!c || !corr) return ccd::fail(CCD_E_INVALID, "ccd_mbpt2: null argument");
    if (c->nprob > 1) return batch_only(c, "ccd_mbpt2");
    if (c->utiles.empty()) {
        *corr = 0.0;
        return CCD_OK;
    }
    UpdArgs a = upd_args(c, nullptr, nullptr, 0.0);
    hipLaunchKernelGGL(k_mbpt2, dim3((unsigned)c->utiles.size()), dim3(kThreads), 0, c->stream, a, c->d_partials);
    int rc = check_launch("k_mbpt2");
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, c->stream, c->d_partials, (int)c->utiles.size(),
                       0.25, c->d_scalar);
    if ((rc = check_launch("k_sum_partials"))) return rc;
    return read_scalar(c, corr);
}

int ccd_mbpt2_noalloc(const ccd_ctx *c, double *d_corr, void *stream) {
    ccd::clear_error();
    if (!c || !d_corr) return ccd::fail(CCD_E_INVALID, "ccd_mbpt2_noalloc: null argument");
    if (c->nprob > 1) return batch_only(c, "ccd_mbpt2_noalloc");
    const int n = (int)c->utiles.size();
    if (n > kNoAllocPartials)
        return ccd::fail(CCD_E_INVALID, "ccd_mbpt2_noalloc: the problem has %d update tiles, the scratch holds %d", n,
                         kNoAllocPartials);
    double *ptr = scratch_ptr(c->device);
    if (!ptr) return ccd::fail(CCD_E_STATE, "ccd_mbpt2_noalloc: no scratch address for device %d", c->device);
    const hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    if (n > 0) {   // every partial is written before it is read: nothing to clear
        UpdArgs a = upd_args(c, nullptr, nullptr, 0.0);
        hipLaunchKernelGGL(k_mbpt2, dim3((unsigned)n), dim3(kThreads), 0, st, a, ptr);
        if (int rc = check_launch("k_mbpt2")) return rc;
    }
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, st, ptr, n, 0.25, d_corr);
    return check_launch("k_sum_partials");
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
    if (c->nprob > 1) return batch_only(c, "ccd_iterate");
    int rc = arm_free_run(c);
    if (rc || (rc = run_n(c, mixing, 1))) return rc;
    double e = 0.0;
    if ((rc = read_scalar(c, &e))) return rc;
    if (std::isnan(e)) mark_stopped(c);   // the loop state stopped this iteration
    if (corr) *corr = e;
    return CCD_OK;
}

int ccd_solve(ccd_ctx *c, double tol, int max_iters, double mixing, double *energies, int cap, int *n_iter,
              int *stop) {
    ccd::clear_error();
    if (!c) return ccd::fail(CCD_E_INVALID, "null context");
    if (c->nprob > 1) return batch_only(c, "ccd_solve");
    if (max_iters < 0) return ccd::fail(CCD_E_INVALID, "max_iters < 0");
    int rc = zero_state(c);
    if (rc || (rc = arm_loop(c, tol, max_iters, true)) || (rc = run_loop(c, mixing, kSolveBatch))) return rc;
    const LoopCtl st = *c->h_ctl;
    if ((rc = copy_log(c, energies, cap, st.count))) return rc;
    if (n_iter) *n_iter = st.count;
    if (stop) *stop = st.reason;
    return end_solve(c);
}

int ccd_set_precision(ccd_ctx *c, int precision) {
    ccd::clear_error();
    if (!c) return ccd::fail(CCD_E_INVALID, "null context");
    if (precision != CCD_PREC_FP64 && precision != CCD_PREC_FP32)
        return ccd::fail(CCD_E_INVALID, "unknown precision %d", precision);
    if (precision == CCD_PREC_FP32 && c->precision != CCD_PREC_FP32)
        return ccd::fail(CCD_E_STATE, "FP32 iterations need a context created with CCD_PREC_FP32");
    c->iter_prec = precision;
    return sync_state(c);
}

int ccd_set_schedule(ccd_ctx *c, int sched) {
    ccd::clear_error();
    if (!c) return ccd::fail(CCD_E_INVALID, "null context");
    if (sched < CCD_SCHED_AUTO || sched > CCD_SCHED_SERIAL_GRAPH)
        return ccd::fail(CCD_E_INVALID, "unknown schedule %d", sched);
    if (sched == CCD_SCHED_AUTO) {
        const int saved = c->iter_prec;
        int rc = tune_schedule(c);
        c->iter_prec = saved;
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(c->stream));
        return CCD_OK;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (int rc = drop_graphs(c)) return rc;
    c->sched = sched;
    return CCD_OK;
}

int ccd_get_ladder_plan(const ccd_ctx *c, ccd_ladder_plan *plan) {
    ccd::clear_error();
    if (!c || !plan) return ccd::fail(CCD_E_INVALID, "ccd_get_ladder_plan: null argument");
    *plan = c->plan;
    return CCD_OK;
}

int ccd_set_ladder_plan(ccd_ctx *c, const ccd_ladder_plan *plan) {
    ccd::clear_error();
    if (!c || !plan) return ccd::fail(CCD_E_INVALID, "ccd_set_ladder_plan: null argument");
    if (int rc = plan_valid(*plan)) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    ccd_ladder_plan p = *plan;
    for (int q = 0; q < 2; q++)   // the measurements stay those of the tuner
        for (int n = 0; n < NTMAX; n++) p.us[q][n] = c->plan.us[q][n];
    p.us_partition[0] = c->plan.us_partition[0];
    p.us_partition[1] = c->plan.us_partition[1];
    c->plan = p;
    plan_ladder_tiles(c);
    int rc = upload_ladder_tiles(c);
    return rc ? rc : drop_graphs(c);
}

int ccd_tune_ladder(ccd_ctx *c, int flags) {
    ccd::clear_error();
    if (!c) return ccd::fail(CCD_E_INVALID, "null context");
    if (flags & ~(CCD_TUNE_KINDS | CCD_TUNE_BF16X3 | CCD_TUNE_RING_BF16X3))
        return ccd::fail(CCD_E_INVALID, "unknown tuner flags %d", flags);
    const int saved = c->iter_prec;
    int rc = tune_ladder(c, flags);
    c->iter_prec = saved;
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int ccd_get_schedule(const ccd_ctx *c, int *sched, double *us_per_iter) {
    ccd::clear_error();
    if (!c) return ccd::fail(CCD_E_INVALID, "null context");
    if (sched) *sched = c->sched;
    if (us_per_iter)
        for (int q = 0; q < 3; q++) us_per_iter[q] = c->sched_us[q];
    return CCD_OK;
}

int ccd_solve_mixed(ccd_ctx *c, double tol, double switch_tol, int max_iters, double mixing, double *energies,
                    int cap, int *n_iter, int *n_fp32, int *stop) {
    ccd::clear_error();
    if (!c) return ccd::fail(CCD_E_INVALID, "null context");
    if (c->nprob > 1) return batch_only(c, "ccd_solve_mixed");
    if (c->precision != CCD_PREC_FP32)
        return ccd::fail(CCD_E_STATE, "ccd_solve_mixed needs a context created with CCD_PREC_FP32");
    if (max_iters < 0) return ccd::fail(CCD_E_INVALID, "max_iters < 0");
    const int saved = c->iter_prec;
    int rc = zero_state(c);
    // FP32 contractions while |dE| > max(tol, switch_tol) ...
    c->iter_prec = CCD_PREC_FP32;
    c->tph_prec = CCD_PREC_FP32;   // T^ph and tas are zero: the same in either type
    c->tas_prec = CCD_PREC_FP32;
    if (!rc) rc = arm_loop(c, std::max(tol, switch_tol), max_iters, true);
    if (!rc) rc = run_loop(c, mixing, kSolveBatch);
    int n32 = rc ? 0 : c->h_ctl->count;
    // ... then FP64 sweeps, continuing the same loop, while |dE| > tol
    if (!rc && c->h_ctl->reason == CCD_STOP_CONVERGED && switch_tol > tol) {
        c->iter_prec = CCD_PREC_FP64;
        rc = arm_loop(c, tol, max_iters, false);
        if (!rc) rc = sync_state(c);
        if (!rc) rc = run_loop(c, mixing, kSolveBatch);
    }
    c->iter_prec = saved;   // the derived state follows at the next iteration (sync_state)
    if (rc) return rc;
    const LoopCtl st = *c->h_ctl;
    if ((rc = copy_log(c, energies, cap, st.count))) return rc;
    if (n_iter) *n_iter = st.count;
    if (n_fp32) *n_fp32 = n32;
    if (stop) *stop = st.reason;
    return end_solve(c);
}

int ccd_iterate_n(ccd_ctx *c, int n, double mixing, double *energies) {
    ccd::clear_error();
    if (!c || n < 0) return ccd::fail(CCD_E_INVALID, "ccd_iterate_n: bad argument");
    if (c->nprob > 1) return batch_only(c, "ccd_iterate_n");
    int rc = arm_free_run(c);
    if (rc || (rc = run_n(c, mixing, n))) return rc;
    HIP_TRY(hipMemcpyAsync(c->h_ctl, c->d_ctl, sizeof(LoopCtl), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->h_ctl->count != n) {   // a NaN energy stops even a free run
        c->cur = (c->cur + c->h_ctl->count - n) & 1;
        mark_stopped(c);
        return ccd::fail(CCD_E_STATE, "iteration %d produced a NaN energy", c->h_ctl->count);
    }
    return copy_log(c, energies, n, n);
}

int ccd_get_t2(ccd_ctx *c, double *t2) {
    ccd::clear_error();
    if (!c || !t2) return ccd::fail(CCD_E_INVALID, "ccd_get_t2: null argument");
    if (int rc = t2_download(c, c->d_t[c->cur], t2, c->d_g)) return rc;
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
    if (int rc0 = arm_free_run(c)) return rc0;
    double *dt = c->d_t[1 - c->cur];  // scratch: the non-current T buffer
    int rc;
    if ((rc = t2_upload(c, t, dt, c->d_g))) return rc;
    rc = launch_antisym(c, dt);
    c->tas_prec = kStale;   // the caller's rows, not those of the current amplitudes
    if (rc || (rc = launch_ladder_x(c, c->stream)) || (rc = launch_ladder_tx(c, c->stream)) ||
        (rc = launch_ladder_v(c, 1, NTMAX, c->stream)))
        return rc;
    if (!c->utiles.empty()) {
        UpdArgs a = upd_args(c, nullptr, nullptr, 0.0);
        if (c->iter_prec == CCD_PREC_FP32)
            hipLaunchKernelGGL(k_expand_d<float>, dim3((unsigned)c->utiles.size()), dim3(kThreads), 0, c->stream, a,
                               c->d_g);
        else
            hipLaunchKernelGGL(k_expand_d<double>, dim3((unsigned)c->utiles.size()), dim3(kThreads), 0, c->stream, a,
                               c->d_g);
        if ((rc = check_launch("k_expand_d"))) return rc;
    }
    if ((rc = t2_download(c, c->d_g, d, c->d_r))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int ccd_ring(ccd_ctx *c, const double *tph, double *r) {
    ccd::clear_error();
    if (!c || !tph || !r) return ccd::fail(CCD_E_INVALID, "ccd_ring: null argument");
    if (int rc0 = arm_free_run(c)) return rc0;
    if (c->iter_prec == CCD_PREC_FP32) {   // T^ph in FP32, as the FP32 iterations hold it
        HIP_TRY(hipMemcpyAsync(c->d_g, tph, sizeof(double) * c->nt2, hipMemcpyHostToDevice, c->stream));
        if (c->nt2 > 0) {
            const unsigned g = (unsigned)std::min<int64_t>((c->nt2 + kThreads - 1) / kThreads, 4096);
            hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(kThreads), 0, c->stream, c->d_g,
                               reinterpret_cast<float *>(c->d_tph), c->nt2);
            if (int rc1 = check_launch("k_to_f32")) return rc1;
        }
    } else {
        HIP_TRY(hipMemcpyAsync(c->d_tph, tph, sizeof(double) * c->nt2, hipMemcpyHostToDevice, c->stream));
    }
    c->tph_prec = kStale;   // the caller's T^ph, not that of the current amplitudes
    int rc = launch_ring(c, c->stream);
    if (rc) return rc;
    if (c->iter_prec == CCD_PREC_FP32) {   // R in FP32 (exactly the FP32 sums)
        std::vector<float> r32((size_t)c->nt2);
        HIP_TRY(hipMemcpyAsync(r32.data(), c->d_r, sizeof(float) * c->nt2, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (int64_t k = 0; k < c->nt2; k++) r[k] = r32[k];
        return CCD_OK;
    }
    HIP_TRY(hipMemcpyAsync(r, c->d_r, sizeof(double) * c->nt2, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CCD_OK;
}

int ccd_remap_ph(ccd_ctx *c, const double *t, double *tph) {
    ccd::clear_error();
    if (!c || !t || !tph) return ccd::fail(CCD_E_INVALID, "ccd_remap_ph: null argument");
    if (int rc0 = arm_free_run(c)) return rc0;
    double *dt = c->d_t[1 - c->cur];
    int rc = t2_upload(c, t, dt, c->d_g);
    if (!rc) rc = launch_transpose(c, true, dt, c->d_tph, c->stream);
    c->tph_prec = kStale;   // the caller's T^ph
    if (rc) return rc;
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
    if (c->nprob > 1) return batch_only(c, "ccd_step");
    int rc = set_state(c, t_old);
    if (rc) return rc;
    if ((rc = run_n(c, mixing, 1))) return rc;
    double e = 0.0;
    if ((rc = read_scalar(c, &e))) return rc;
    if ((rc = t2_download(c, c->d_t[c->cur], t_new, c->d_g))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (std::isnan(e)) mark_stopped(c);
    if (corr) *corr = e;
    return CCD_OK;
}

}  // extern "C"

namespace {

// Average device time of each launch group over `iters` iterations run in
// order on the main stream, each group launched `reps` times back to back
// (acc[CCD_KT_COUNT] = the whole iteration, acc[k] per launch of group k).
int time_groups(ccd_ctx *c, int iters, double mixing, int reps, double (&acc)[CCD_KT_COUNT + 1]) {
    if (int rc0 = arm_free_run(c)) return rc0;
    hipEvent_t ev[CCD_KT_COUNT + 1];
    for (int k = 0; k <= CCD_KT_COUNT; k++) HIP_TRY(hipEventCreate(&ev[k]));
    for (double &a : acc) a = 0.0;
    int rc = CCD_OK;
    for (int it = 0; it < iters && !rc; it++) {
        rc = run_iteration(c, mixing, ev, reps);
        if (rc) break;
        hipError_t e = hipEventSynchronize(ev[CCD_KT_COUNT]);
        if (e != hipSuccess) {
            rc = ccd::fail(CCD_E_HIP, "iteration failed: %s", hipGetErrorString(e));
            break;
        }
        for (int k = 0; k < CCD_KT_COUNT; k++) {
            float f = 0.f;
            (void)hipEventElapsedTime(&f, ev[k], ev[k + 1]);
            acc[k] += f / reps;
        }
        float f = 0.f;
        (void)hipEventElapsedTime(&f, ev[0], ev[CCD_KT_COUNT]);
        acc[CCD_KT_COUNT] += f;
    }
    for (int k = 0; k <= CCD_KT_COUNT; k++) (void)hipEventDestroy(ev[k]);
    if (rc) return rc;
    for (double &a : acc) a /= iters;
    return CCD_OK;
}

}  // namespace

extern "C" {

int ccd_time_phases(ccd_ctx *c, int iters, double mixing, double *ms) {
    ccd::clear_error();
    if (!c || !ms || iters <= 0) return ccd::fail(CCD_E_INVALID, "ccd_time_phases: bad argument");
    double g[CCD_KT_COUNT + 1];
    if (int rc = time_groups(c, iters, mixing, 1, g)) return rc;
    ms[0] = g[CCD_KT_RING] + g[CCD_KT_REMAP_R];
    ms[1] = g[CCD_KT_X];
    ms[2] = g[CCD_KT_TX] + g[CCD_KT_LADDER1] + g[CCD_KT_LADDER2] + g[CCD_KT_LADDER34] + g[CCD_KT_LADDER58];
    ms[3] = g[CCD_KT_UPDATE];
    ms[4] = g[CCD_KT_REMAP_T] + g[CCD_KT_REDUCE];
    ms[5] = g[CCD_KT_COUNT];
    return CCD_OK;
}

int ccd_time_kernels(ccd_ctx *c, int iters, int reps, double mixing, double *ms, int n) {
    ccd::clear_error();
    if (!c || !ms || iters <= 0 || reps <= 0 || n < 0)
        return ccd::fail(CCD_E_INVALID, "ccd_time_kernels: bad argument");
    double g[CCD_KT_COUNT + 1];
    if (int rc = time_groups(c, iters, mixing, reps, g)) return rc;
    for (int k = 0; k < n && k <= CCD_KT_COUNT; k++) ms[k] = g[k];
    return CCD_OK;
}

}  // extern "C"
