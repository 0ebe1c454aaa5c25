/*
 * ccd_amd.h -- C ABI of the MI355X-native coupled-cluster-doubles iteration.
 *
 * Two halves:
 *
 *  1. The device context (ccd_ctx_*, ccd_iterate, ccd_solve, per-diagram
 *     entry points).  This is the drop-in for the reference's hot path:
 *     the diagram functions of src/calc.cpp, MBPT2Corr of
 *     src/MBPT2Corr.cpp, the Chain remaps of src/Chain.hpp:710-802 and the
 *     iteration loop of src/main.cpp:491-682.  It takes a plain descriptor
 *     of the channel-blocked problem (ccd_problem) that a caller fills from
 *     the reference's SPBasis arrays and Chain<double> blocks.
 *
 *  2. The host problem model (ccd_model_*), a behaviour-identical restatement
 *     of the reference's basis setup (src/PairingSPBasis.cpp,
 *     src/ElectronGasSPBasis.cpp, src/SPBasis.cpp, src/balance.cpp at one
 *     rank, src/load.cpp), so the library runs the reference's input
 *     format without the reference present.  It fills a ccd_problem.
 *
 * Conventions: every function returns 0 on success and a negative
 * CCD_E_* code on failure (ccd_last_error() gives the message of the
 * calling thread's last failure); nothing throws across the ABI.  Host
 * buffers belong to the caller; the context owns its device memory and one
 * HIP stream.  A context is used from one host thread at a time.
 *
 * Layouts are the reference's Chain<double> layouts (src/Chain.hpp:373-426):
 * channel blocks back to back in the caller's channel order, each block
 * row-major.  "pphh": block c is pp_dim[c] x hh_dim[c], element
 * [(a,b),(i,j)] = t_ij^ab.  "phhp": mod block m is ph_dim[m] x hp_dim[m],
 * element [(a,i),(j,b)] = t_ij^ab.  Pair lists hold ordered pairs (p,q),
 * two int32 per pair, in the row/column order of the blocks.
 */
#ifndef CCD_AMD_H
#define CCD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCD_OK 0
#define CCD_E_INVALID (-1)   /* bad argument or inconsistent descriptor */
#define CCD_E_HIP (-2)       /* HIP runtime failure (no device, launch error) */
#define CCD_E_NOMEM (-3)     /* host or device allocation failed */
#define CCD_E_STATE (-4)     /* call not valid in the context's state */

/* Arithmetic of the amplitude contractions. */
#define CCD_PREC_FP64 0      /* everything FP64 (the parity path) */
#define CCD_PREC_FP32 1      /* mixed: FP32 MFMA contractions (pppp ladder and ring,
                                FP32 accumulation), FP64 hhhh term, update,
                                denominators, one-body sums and energy; the
                                context can also run FP64 iterations */
/* OR-ed into ccd_ctx_create's precision: the matrix-free mode of the reference
 * (-m 1, the calcFree* twins of src/calc.cpp:45-504 driven by
 * src/main.cpp:521-588).  V^pppp is never stored: the pppp ladder rebuilds
 * each element from the generator as it goes, bit-identical to the stored
 * mode.  Needs gen_kind != CCD_GEN_NONE. */
#define CCD_MATRIX_FREE 0x100

/* Interaction source for a ccd_problem. */
#define CCD_GEN_NONE 0           /* V blocks supplied by the caller (v_* below) */
#define CCD_GEN_PAIRING 1        /* generate <pq||rs> of the pairing model on device */
#define CCD_GEN_ELECTRON_GAS 2   /* generate the antisymmetrised Coulomb element on device */
#define CCD_GEN_INF_MATTER 3     /* generate the antisymmetrised Minnesota element of pure
                                    neutron matter on device */

/* Solve stop reasons (ccd_solve). */
#define CCD_STOP_CONVERGED 0
#define CCD_STOP_MAX_ITERS 1     /* main.cpp:512-519: halted after iteration max_iters */
#define CCD_STOP_NAN 2           /* |dE| > tol was false because dE is NaN (main.cpp:511) */

typedef struct ccd_problem {
    int32_t n_sp;           /* SPBasis::nSpstates */
    int32_t n_holes;        /* SPBasis::nParticles; states 0..n_holes-1 are holes */
    const double *eps;      /* [n_sp] normal-ordered s.p. energies (SPBasis::spEnergy
                               after rotateSpEnergiesToNormalOrdered) */
    int32_t n_chan;         /* direct channels (SPBasis::nChannels) */
    const int32_t *pp_dim;  /* [n_chan] chanDims[c].ppDim */
    const int32_t *hh_dim;  /* [n_chan] chanDims[c].hhDim */
    const int32_t *pp_pairs;/* chanMaps[c].ppMap, concatenated over c */
    const int32_t *hh_pairs;/* chanMaps[c].hhMap, concatenated over c */
    int32_t n_mod;          /* mod channels (also SPBasis::nChannels in the reference) */
    const int32_t *ph_dim;  /* [n_mod] chanModDims[m].phDim */
    const int32_t *hp_dim;  /* [n_mod] chanModDims[m].hpDim */
    const int32_t *ph_pairs;/* chanModMaps[m].phMap, concatenated */
    const int32_t *hp_pairs;/* chanModMaps[m].hpMap, concatenated */
    /* Caller-supplied interaction (gen_kind == CCD_GEN_NONE), FP64, reference
     * Chain layouts over all channels: vnn_hhhh (hh x hh), vnn_hhpp (hh x pp),
     * vnn_pppp (pp x pp), vnn_hhpp_hpph_mod (hp x ph), vnn_hpph_hphp_mod
     * (hp x hp) -- see src/load.cpp:112-318 and src/main.cpp:293-366. */
    const double *v_hhhh;
    const double *v_hhpp;
    const double *v_pppp;
    const double *v_modA;
    const double *v_modB;
    /* On-device generation (gen_kind != CCD_GEN_NONE). */
    int32_t gen_kind;
    const int32_t *qn;      /* [n_sp*5] electron gas: nx ny nz sz tz; pairing: level spin - - - */
    double gen_g;           /* pairing strength g; infinite matter: the box length L */
    int32_t gen_nmax;       /* electron gas: max |n_x| of the basis */
    const double *gen_table;/* electron gas: direct Coulomb term per momentum transfer,
                               (4*gen_nmax+1)^3 entries, index
                               ((dx+2m)*(4m+1) + (dy+2m))*(4m+1) + (dz+2m), m = gen_nmax;
                               infinite matter: the 3 Minnesota prefactors
                               +-V_0X / L^3 * (pi / kappa_X)^1.5 for X = R, T, S
                               (src/InfMatterSPBasis.cpp:412-414 up to the Gaussian) */
} ccd_problem;

typedef struct ccd_ctx ccd_ctx;
typedef struct ccd_model ccd_model;

const char *ccd_last_error(void);

/* ---------------------------------------------------------------- device */

/* Validate the descriptor, build the device tables and upload (or generate)
 * the interaction on the current HIP device.  T2 starts at zero.
 * precision: CCD_PREC_*, optionally | CCD_MATRIX_FREE. */
int ccd_ctx_create(const ccd_problem *prob, int precision, ccd_ctx **out);
int ccd_ctx_destroy(ccd_ctx *ctx);

/* Sizes: elements of the pphh (= phhp) layout, bytes of device memory held. */
int ccd_ctx_info(const ccd_ctx *ctx, int64_t *n_t2, int64_t *device_bytes);

/* Arithmetic of the contractions of the following iterations: CCD_PREC_FP64
 * (always available) or CCD_PREC_FP32 (contexts created with CCD_PREC_FP32;
 * their initial setting).  No counterpart in the reference, which is FP64
 * only. */
int ccd_set_precision(ccd_ctx *ctx, int precision);

/* How the iterations are launched (ccd_set_schedule).  No counterpart in the
 * reference.  Every schedule runs the same kernels in an order that keeps
 * their dependencies, and every reduction is in a fixed order, so all give
 * the same results bit for bit. */
#define CCD_SCHED_AUTO 0           /* the fastest of the three below, timed on the
                                      context's own problem */
#define CCD_SCHED_STREAMS 1        /* independent phases on four streams, launched
                                      iteration by iteration (the initial one) */
#define CCD_SCHED_SERIAL 2         /* every launch in order on one stream */
#define CCD_SCHED_SERIAL_GRAPH 3   /* the one-stream iterations captured as a HIP
                                      graph of 4 iterations and replayed */

/* Choose the schedule of the following iterations.  CCD_SCHED_AUTO times the
 * three on this problem (a few free iterations) and resets T2 to zero; the
 * others keep the amplitudes. */
int ccd_set_schedule(ccd_ctx *ctx, int sched);
/* The schedule in use (never CCD_SCHED_AUTO) and, if us_per_iter is not NULL,
 * the tuner's last measurements (0 before any): microseconds per iteration
 * of schedules 1..3 in us_per_iter[0..2]. */
int ccd_get_schedule(const ccd_ctx *ctx, int *sched, double *us_per_iter);

/* The launch plan of the pppp ladder (the ladder of calcChannelT2Diff_pphh,
 * src/calc.cpp:188-201; its counterpart in the reference is the static
 * channel distribution of src/balance.cpp:91-144 and src/main.cpp:94-115).
 * No counterpart in the reference.  Channels are bucketed by size into 8
 * classes: class n holds the channels with 16 (n-1) < hh <= 16 n.  Per
 * class the plan fixes the kernel (kind, per arithmetic of the iterations),
 * how many persistent blocks per compute unit run its tiles, and how its
 * tiles are dealt to them; `partition` fixes which classes share one grouped
 * launch.  The block count, the assignment and the partition change no
 * result (tiles write disjoint outputs); the kernel kind does (another
 * summation order, or split products). */
#define CCD_LADDER_CLASSES 8
#define CCD_LK_DEFAULT 0     /* FP64: v_mfma_f64_16x16x4f64; FP32: the one-row-per-lane
                                v_mfma_f32_4x4x1f32 kernel for class 1, v_mfma_f32_16x16x4f32
                                for the others */
#define CCD_LK_MFMA16 1      /* v_mfma_{f64,f32}_16x16x4 in every class */
#define CCD_LK_BF16X3 2      /* FP32 iterations only: BF16x3 split products (x = x_hi + x_lo
                                in bfloat16, a b ~ a_lo b_hi + a_hi b_lo + a_hi b_hi on
                                v_mfma_f32_16x16x32_bf16, FP32 accumulation; about 2^-16
                                relative per product against 2^-24 for FP32) */
#define CCD_LK_BF16X3SYM 3   /* FP32 iterations, class 1 only: BF16x3 split products on the
                                symmetry of V^pppp (<ab||cd> = <cd||ab>): each element with
                                c,d at or past the row tile of a,b is built once and used
                                for both rows, about 0.55 of the elements of the others */
#define CCD_LK_BF16X3W 4     /* FP32 iterations, class 1 only: the kernel of CCD_LK_BF16X3SYM
                                without the symmetry (every element built; waves split k,
                                four rows per lane) */
#define CCD_RK_FP32 0        /* ring: v_mfma_f32_16x16x4f32 (the start) */
#define CCD_RK_BF16X3 1      /* ring: BF16x3 split products on v_mfma_f32_16x16x16_bf16 */
#define CCD_LA_MODEL 0       /* tile assignment: the one a model of the CU loads prefers
                                (a matrix-free batch: CCD_LA_RUNS) */
#define CCD_LA_CHANNEL 1     /*   round robin in channel order */
#define CCD_LA_LONGEST 2     /*   longest tile first to the least loaded block */
#define CCD_LA_RUNS 3        /*   contiguous runs of about equal cost, in channel order */
#define CCD_LA_XCD 4         /*   channels (the largest cut into pieces of at most 1/32 of the
                                class) to the 8 XCDs, least loaded first, then longest tile
                                first over each XCD's blocks: the tiles of a channel, which
                                read the same packed T rows, share one L2 */
#define CCD_LP_SPLIT 0       /* launches: classes 1, 2, 3-4 and 5-8 (in the four-stream
                                schedule classes 1-2 and 3-8 run on two streams) */
#define CCD_LP_ONE 1         /* one grouped launch of every class */
typedef struct ccd_ladder_plan {
    int32_t kind[2][CCD_LADDER_CLASSES];   /* [CCD_PREC_FP64 / CCD_PREC_FP32][class - 1]: CCD_LK_* */
    int32_t blocks_per_cu[CCD_LADDER_CLASSES];  /* 0: one block per tile; k > 0: k blocks
                                                   per compute unit, each running a run of tiles */
    int32_t assign[CCD_LADDER_CLASSES];    /* CCD_LA_* */
    int32_t partition;                     /* CCD_LP_* */
    double us[2][CCD_LADDER_CLASSES];      /* set by ccd_tune_ladder: microseconds per launch of
                                              each class as chosen (0: not measured) */
    double us_partition[2];                /* ccd_tune_ladder: microseconds per iteration of
                                              CCD_LP_SPLIT and CCD_LP_ONE (0: not measured) */
    int32_t ring_fp32;                     /* the ring of FP32 iterations (calcChannelT2Diff_phhp,
                                              src/calc.cpp:264-313) in its mod channels of 1-5
                                              column tiles: CCD_RK_* */
} ccd_ladder_plan;
/* The plan in use.  A context starts with CCD_LK_DEFAULT everywhere, 3
 * blocks per compute unit (generated V) or one block per tile (stored V),
 * CCD_LA_MODEL and CCD_LP_SPLIT. */
int ccd_get_ladder_plan(const ccd_ctx *ctx, ccd_ladder_plan *plan);
/* Use `plan` from the next iteration on (its us fields are ignored).  Fails
 * with CCD_E_INVALID on an unknown kind, assignment or partition, a
 * CCD_LK_BF16X3, CCD_LK_BF16X3SYM or CCD_LK_BF16X3W entry for FP64
 * iterations, a CCD_LK_BF16X3SYM or CCD_LK_BF16X3W entry outside class 1,
 * a block count outside 0..16, or an unknown ring kind. */
int ccd_set_ladder_plan(ccd_ctx *ctx, const ccd_ladder_plan *plan);
/* Tune the plan on this problem, for the arithmetic of the following
 * iterations.  Per class: the kernel kinds that `flags` allow, then the
 * blocks per compute unit (1, 2, 3, 4, 6; stored V also one block per tile),
 * then the assignment, each timed with HIP events over repeated launches of
 * that class alone (us).  Then whole iterations in the context's schedule
 * decide among the plan in use, that per-class result and the plan in use
 * with each allowed kind in every class (BF16x3 also with class 1 on
 * CCD_LK_BF16X3SYM or CCD_LK_BF16X3W), with CCD_TUNE_RING_BF16X3 each also
 * with the BF16x3 ring, each under both partitions (us_partition: the
 * chosen plan's).  Resets T2 to zero. */
#define CCD_TUNE_KINDS 1     /* also try CCD_LK_MFMA16 where it differs from the default */
#define CCD_TUNE_BF16X3 2    /* also try CCD_LK_BF16X3 and, for class 1, CCD_LK_BF16X3SYM and
                                CCD_LK_BF16X3W (FP32 iterations) */
#define CCD_TUNE_RING_BF16X3 4  /* also try the ring on CCD_RK_BF16X3 (FP32 iterations).  Its
                                rounding moves the fixed point of the FP32 phase of a mixed
                                solve further: at N=54 the phase then stops after 24
                                iterations and needs 10 FP64 ones instead of 26 + 5. */
int ccd_tune_ladder(ccd_ctx *ctx, int flags);

/* MBPT2 correlation energy, 1/4 sum v^2/(e_i+e_j-e_a-e_b)
 * (replaces MBPT2Corr, src/MBPT2Corr.cpp:6-38). */
int ccd_mbpt2(ccd_ctx *ctx, double *corr);

/* ccd_mbpt2 that allocates nothing and does not synchronise.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory or
 * wait on the host (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  The MBPT2 energy of the context's problem is
 * written to d_corr, one double in device memory, by kernels queued on
 * `stream` (a hipStream_t of the context's device; NULL = the context's own
 * stream); it is there once that stream has reached them.  The per-tile
 * partial sums live in a fixed per-GPU scratch array of 65,536 doubles
 * (512 KB), shared by every call on that GPU: calls that overlap on one GPU
 * race.  Fails with CCD_E_INVALID for problems of more than 65,536 update
 * tiles (the electron gas at N=66, n^2 <= 52, has 15,858) and CCD_E_STATE on
 * a batch.  Same kernels and summation order as ccd_mbpt2: the same value
 * bit for bit. */
int ccd_mbpt2_noalloc(const ccd_ctx *ctx, double *d_corr, void *stream);

/* T2 = 0 and the one-body intermediates = 0 (src/main.cpp:497-503). */
int ccd_reset(ccd_ctx *ctx);

/* One Jacobi step on the device (the loop body of src/main.cpp:589-670):
 * both ladders and the hhhh quadratic term (calcT2Diff_pphh), the ring
 * (calcT2Diff_phhp), the one-body terms (calcT2Diff_h2pp_h1,
 * calcT2Diff_p2_p1hh), all remaps, the update / denominator / mixing
 * (calcT2_pphh) and the correlation energy of the new amplitudes. */
int ccd_iterate(ccd_ctx *ctx, double mixing, double *corr);

/* n iterations queued back to back with no host round trip in between;
 * energies[k] (if not NULL) receives the energy after the k-th.  Fails with
 * CCD_E_STATE if an energy is NaN (the iterations after it do nothing). */
int ccd_iterate_n(ccd_ctx *ctx, int n, double mixing, double *energies);

/* The reference protocol (src/main.cpp:491-519) from T2 = 0: loop while
 * |dE| > tol, dE starting at 1 and E_prev at 100, halt once more than
 * max_iters iterations have run (the reference uses max_iters = 100).
 * energies[k] receives the energy after iteration k for k < min(cap, 1024).
 * The stop rule runs on the device after every iteration; the host queues
 * iterations in small batches and reads the outcome between them. */
int ccd_solve(ccd_ctx *ctx, double tol, int max_iters, double mixing,
              double *energies, int cap, int *n_iter, int *stop);

/* The same protocol with mixed precision (contexts created with
 * CCD_PREC_FP32): FP32 contractions while |dE| > switch_tol, then FP64
 * iterations (defect-correction sweeps back to the FP64 fixed point) until
 * |dE| <= tol.  n_fp32 receives the number of FP32 iterations.  The
 * context's precision setting is left as it was.  No counterpart in the
 * reference. */
int ccd_solve_mixed(ccd_ctx *ctx, double tol, double switch_tol, int max_iters,
                    double mixing, double *energies, int cap, int *n_iter,
                    int *n_fp32, int *stop);

/* Copy the current amplitudes out / in (pphh layout, host memory). */
int ccd_get_t2(ccd_ctx *ctx, double *t2);
int ccd_set_t2(ccd_ctx *ctx, const double *t2);

/* Per-diagram entry points on caller amplitudes (host arrays), for parity
 * tests of each piece against calc.cpp.  ccd_ladder, ccd_ring and
 * ccd_remap_ph leave the current amplitudes in place (work arrays they
 * overwrite are rebuilt from them before the next iteration); ccd_onebody
 * and ccd_step make t (t_new) the current amplitudes.  ccd_ladder computes the ladder of the
 * part of t antisymmetric in (a,b) (V^pppp is stored for pairs a<b only);
 * the reference's amplitudes are always antisymmetric, so on them it is the
 * reference's calcT2Diff_pphh.
 *   ccd_ladder:   d = calcT2Diff_pphh(t)          (calc.cpp:132-231), pphh
 *   ccd_ring:     r = calcT2Diff_phhp(tph)        (calc.cpp:264-339), phhp
 *   ccd_remap_ph: tph = mapToDistributed_ai_jb(t) (Chain.hpp:757-802)
 *   ccd_onebody:  chi[n_holes], psi[n_sp-n_holes], the `value` scalars of
 *                 calcChannelT2Diff_h2pp_h1 / _p2_p1hh (calc.cpp:81-84, 15-20)
 *   ccd_step:     t_new = one full iteration from t_old, returns its energy */
int ccd_ladder(ccd_ctx *ctx, const double *t, double *d);
int ccd_ring(ccd_ctx *ctx, const double *tph, double *r);
int ccd_remap_ph(ccd_ctx *ctx, const double *t, double *tph);
int ccd_onebody(ccd_ctx *ctx, const double *t, double *chi, double *psi);
int ccd_step(ccd_ctx *ctx, const double *t_old, double *t_new, double mixing,
             double *corr);

/* Run `iters` iterations from the current amplitudes and report the average
 * device time of each phase in milliseconds, measured with HIP events on
 * the context's stream: ms[0] ring, ms[1] ladder X (hhhh + quadratic),
 * ms[2] ladder D (pppp), ms[3] update + remap, ms[4] reductions,
 * ms[5] whole iteration.  ms must hold 6 doubles. */
int ccd_time_phases(ccd_ctx *ctx, int iters, double mixing, double *ms);

/* Launch groups of one iteration, in the order they run in the timed
 * iterations of ccd_time_phases / ccd_time_kernels (all on the context's
 * stream). */
#define CCD_KT_RING 0       /* calcT2Diff_phhp: X = V^B + 1/2 V^A T^ph, R = T^ph X */
#define CCD_KT_REMAP_R 1    /* R to the pphh layout (the 4 gathers of main.cpp:636-639) */
#define CCD_KT_X 2          /* X = V^hhhh + 1/2 V^hhpp T (calc.cpp:171-187) */
#define CCD_KT_TX 3         /* 1/2 T X into the ladder output */
#define CCD_KT_LADDER1 4    /* 1/2 V^pppp T, channels with hh <= 16 */
#define CCD_KT_LADDER2 5    /*   17 <= hh <= 32 */
#define CCD_KT_LADDER34 6   /*   33 <= hh <= 64 */
#define CCD_KT_LADDER58 7   /*   65 <= hh <= 128 */
#define CCD_KT_UPDATE 8     /* calcT2_pphh: assembly, denominators, mixing, W sums */
#define CCD_KT_REMAP_T 9    /* T to T^ph (mapToDistributed, Chain.hpp:757-802) */
#define CCD_KT_REDUCE 10    /* chi, psi, the energy and the stop rule */
#define CCD_KT_COUNT 11

/* Like ccd_time_phases, per launch group: `iters` iterations with the groups
 * in order on the context's stream, each group launched `reps` times back to
 * back (the groups rewrite the same outputs from the same inputs, so a
 * repetition changes no result; the reductions also count the extra runs in
 * the loop state of the free run), HIP events around each group.  ms[k]
 * (k < n): the average time of one launch of group k in milliseconds;
 * ms[CCD_KT_COUNT] (if n > CCD_KT_COUNT): the average time of a whole
 * timed iteration.  No counterpart in the reference. */
int ccd_time_kernels(ccd_ctx *ctx, int iters, int reps, double mixing, double *ms, int n);

/* Algorithmic work of one iteration in the reference's ordered-pair algebra:
 * flops F = sum_c (2hh^2 pp + 2pp hh^2 + 2pp^2 hh) + sum_m (2hp^2 ph + 2ph hp^2)
 * and the pppp-ladder part of it; bytes one pass of the packed pppp ladder
 * moves: V^pppp for pairs a<b, c<d (none in the matrix-free mode), the
 * antisymmetrised amplitudes and X read once, D_p written once (V^pppp, the
 * amplitude rows and D_p in the arithmetic type of the iterations, X in
 * FP64). */
int ccd_work(const ccd_ctx *ctx, double *flops, double *flops_pppp,
             double *bytes_vpppp);

/* Algorithmic work of each launch group (CCD_KT_* order, k < n) of one
 * iteration in the arithmetic of the next iterations: flops[k] the flops of
 * its products on the packed pairs a<b the device uses (the reference's
 * ordered pairs do 4x the pppp ladder's flops), bytes[k] the bytes its
 * operands and results take when each is read or written once (contraction
 * operands and outputs in the iteration's type, amplitudes, V^hhpp and X in
 * FP64; no index tables), elements[k] the V^pppp elements the matrix-free
 * ladder generates (0 otherwise).  Any pointer may be NULL.  No counterpart
 * in the reference. */
int ccd_kernel_work(const ccd_ctx *ctx, double *flops, double *bytes,
                    double *elements, int n);

/* ---------------------------------------------------------------- batches
 *
 * Many independent solves in one context -- the parameter sweep, which the
 * reference runs as one ccd.exe process per basis (src/main.cpp is one
 * solve).  No counterpart in the reference.  The problems are merged into
 * one channel-blocked problem: states renumbered holes first (the holes of
 * problem 0, of problem 1, ..., then the particles in the same order),
 * channels and their chains back to back in problem order.  Every kernel of
 * an iteration then covers all solves in one launch.  Each solve keeps its
 * own loop state (E_prev, the stop rule, its energy log); a solve that has
 * stopped keeps its amplitudes while the others go on, and the iteration
 * ends once all have stopped.  Each solve's energies and amplitudes are
 * bit-identical to those of a context of its own.  The problems must all
 * generate V on the device (gen_kind != CCD_GEN_NONE; pairing and electron
 * gas may be mixed) or all supply it.  ccd_get_t2 / ccd_set_t2 and the
 * per-diagram entry points work on the merged layout; ccd_mbpt2,
 * ccd_iterate, ccd_iterate_n, ccd_solve, ccd_solve_mixed and ccd_step fail
 * with CCD_E_STATE on a batch of more than one solve. */
int ccd_batch_create(const ccd_problem *const *probs, int n, int precision,
                     ccd_ctx **out);
/* Number of solves; t2_offset (if not NULL, n+1 entries): solve q's
 * amplitudes are elements [t2_offset[q], t2_offset[q+1]) of ccd_get_t2. */
int ccd_batch_info(const ccd_ctx *ctx, int *n_problems, int64_t *t2_offset);
/* MBPT2 of every solve (corr[n]). */
int ccd_batch_mbpt2(ccd_ctx *ctx, double *corr);
/* ccd_solve for every solve at once: energies[q*cap + k], n_iter[q], stop[q]. */
int ccd_batch_solve(ccd_ctx *ctx, double tol, int max_iters, double mixing,
                    double *energies, int cap, int *n_iter, int *stop);
/* ccd_solve_mixed for every solve at once (same layout, n_fp32[q]); the
 * FP32 phase lasts until every solve has stopped in it. */
int ccd_batch_solve_mixed(ccd_ctx *ctx, double tol, double switch_tol,
                          int max_iters, double mixing, double *energies,
                          int cap, int *n_iter, int *n_fp32, int *stop);
/* ccd_iterate_n for every solve at once: energies[q*n + k].  A solve whose
 * energy turns NaN stops there (its later entries are NaN) while the others
 * go on; this is not an error for a batch. */
int ccd_batch_iterate_n(ccd_ctx *ctx, int n, double mixing, double *energies);

/* ---------------------------------------------------------------- model */

/* Pairing model, reference PairingSPBasis(0, xi, g, nPairStates, nParticles). */
int ccd_model_pairing(int n_pair_states, int n_particles, double g, double xi,
                      ccd_model **out);
/* Electron gas, reference ElectronGasSPBasis(2, r_s, 1, nShells, nParticleShells). */
int ccd_model_electron_gas(double r_s, int n_shells, int n_particle_shells,
                           ccd_model **out);
/* Infinite nuclear matter with the Minnesota potential, reference
 * InfMatterSPBasis(1, density, tzMax, nShells, nParticleShells); pure neutron
 * matter (tz_max = 1) only.  The descriptor generates V on the device
 * (CCD_GEN_INF_MATTER; the device exp can differ from the host's in the last
 * bit, so its elements match the host V chains to about 1e-16 relative). */
int ccd_model_inf_matter(double density, int tz_max, int n_shells,
                         int n_particle_shells, ccd_model **out);
int ccd_model_destroy(ccd_model *m);
/* Reference energy (SPBasis::referenceEnergy, before normal ordering). */
int ccd_model_eref(const ccd_model *m, double *eref);
int ccd_model_sizes(const ccd_model *m, int32_t *n_sp, int32_t *n_holes,
                    int32_t *n_chan, int64_t *n_t2);
/* Descriptor with on-device generation; valid until the model is destroyed. */
int ccd_model_problem(ccd_model *m, const ccd_problem **out);
/* Build the host FP64 V chains (the reference's loadVnn_*) and return a
 * descriptor that carries them instead of a generator. */
int ccd_model_problem_host_v(ccd_model *m, const ccd_problem **out);

#ifdef __cplusplus
}
#endif
#endif
