/*
 * ccd_oracle.c -- CPU restatement of the reference CCD iteration.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker).  See ccd_oracle.h for the
 * list of reference files this follows and for how it is pinned.
 *
 * Reference file:line citations are relative to the reference tree root.
 */
#include "ccd_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct { int32_t p, q; } opair;
typedef struct { int32_t p, q, r; } otriple;

struct orc {
    int model;               /* 0 pairing, 1 infinite matter, 2 electron gas (reference basisIndicator) */
    size_t nsp, nh;
    int (*qn)[5];            /* HEG: nx ny nz sz tz; pairing: level spin */
    double *eps;             /* normal-ordered after setup */
    double eref;
    /* electron gas */
    double r_s, L, e_sq;
    int nmax;
    size_t emax, efermi;
    /* pairing */
    double g, xi;
    /* infinite matter (pure neutron matter) */
    double density, inf_L;
    /* channels, permuted order (balance.cpp ffdBalance at one rank) */
    size_t nch;
    size_t *ch_raw, *ch_inv;
    size_t *pp, *hh, *mph, *mhp;
    opair **ppl, **hhl, **mphl, **mhpl;
    size_t *hpp_dim, *phh_dim;
    otriple **hppl, **phhl;
    /* chain offsets */
    size_t *off_t;           /* pphh / hhpp blocks */
    size_t *off_hhhh, *off_pppp;
    size_t *off_ph, *off_mb; /* phhp (and V^A) blocks, V^B blocks */
    size_t *off_hpp, *off_phh;
    size_t nt2, nhhhh, npppp, nmb;
    /* V chains */
    double *v_hhhh, *v_hhpp, *v_pppp, *v_phh_p, *v_h_hpp, *v_modA, *v_modB;
    /* T2 maps (flat index into the source chain) */
    size_t *m_iab_j, *m_jab_i, *m_a_bij, *m_b_aij;
    size_t *m_ai_jb, *m_aj_ib, *m_bi_ja, *m_bj_ia;
    /* work chains */
    double *t2, *t2old, *d, *r, *tph, *s_jab_i, *s_iab_j, *s_a_bij, *s_b_aij;
    double *diffh, *diffp;
    double *gat[8];
    double *chan_e;
    int loaded;
    int free_pppp;           /* V^pppp rebuilt per channel inside the ladder (-m 1) */
};

/* ------------------------------------------------------------------ utils */

/* Legendre three-square test and integer sqrt ceiling (src/util.cpp:1-19). */
static int three_squares(unsigned n) {
    if (n == 0) return 1;
    while (n % 4 == 0) n /= 4;
    return (n % 8) != 7;
}
static unsigned sqrt_ceil(unsigned x) {
    unsigned sq = 1, delta = 3;
    while (sq < x) { sq += delta; delta += 2; }
    return (delta - 1) / 2;
}

void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

/* Row-major C = alpha*A*B + beta*C (the cblas_dgemm calls of src/calc.cpp
 * are all RowMajor/NoTrans with lda=K, ldb=N, ldc=N). */
static void dgemm_rm(size_t M, size_t N, size_t K, double alpha, const double *A,
                     const double *B, double beta, double *C) {
    for (size_t i = 0; i < M; i++) {
        double *c = C + i * N;
        if (beta == 0.0) {
            for (size_t j = 0; j < N; j++) c[j] = 0.0;
        } else if (beta != 1.0) {
            for (size_t j = 0; j < N; j++) c[j] *= beta;
        }
        for (size_t k = 0; k < K; k++) {
            double a = alpha * A[i * K + k];
            const double *b = B + k * N;
            for (size_t j = 0; j < N; j++) c[j] += a * b[j];
        }
    }
}

/* ------------------------------------------------------------------ TBMEs */

/* Electron gas direct term <pq|v|rs> (src/ElectronGasSPBasis.cpp:424-471). */
static double heg_direct(const orc_t *o, size_t p, size_t q, size_t r, size_t s) {
    const int *qi = o->qn[p], *qj = o->qn[q], *qk = o->qn[r], *ql = o->qn[s];
    if (qi[3] == qk[3] && qj[3] == ql[3]) {
        if (qi[0] + qj[0] == qk[0] + ql[0] && qi[1] + qj[1] == qk[1] + ql[1] &&
            qi[2] + qj[2] == qk[2] + ql[2]) {
            double prefactor = o->e_sq / (o->L * o->L * o->L);
            double relMom = 0.0, mu = 0.0;
            for (int d = 0; d < 3; d++) {
                double term = 2 * M_PI * (qi[d] - qk[d]) / o->L;
                relMom += term * term;
            }
            if (relMom < 1.e-5) return 0.0;
            return prefactor * 4 * M_PI / (relMom + mu * mu);
        }
    }
    return 0.0;
}

/* Pairing antisymmetrised <pq||rs> (src/PairingSPBasis.cpp:127-141). */
static double pairing_tbme(const orc_t *o, size_t p, size_t q, size_t r, size_t s) {
    const int *a = o->qn[p], *b = o->qn[q], *c = o->qn[r], *d = o->qn[s];
    double v = 0.0;
    if (a[0] == b[0] && c[0] == d[0] && a[1] == -b[1] && -c[1] == d[1]) {
        v = -0.5 * o->g;
        if (a[1] == -c[1]) v = -v;
    }
    return v;
}

static int kron_del(int i, int j) { return i != j ? 0 : 1; }          /* InfMatterSPBasis.cpp:461-466 */
static int spin_exchange(int i, int j, int k, int l) {                  /* :452-458 */
    return (i == l && j == k) ? 1 : 0;
}

/* Minnesota potential <pq|v|rs>, not antisymmetrised
 * (src/InfMatterSPBasis.cpp:361-449), same operation order. */
static double minnesota_direct(const orc_t *o, size_t p, size_t q, size_t r, size_t s) {
    const int *qi = o->qn[p], *qj = o->qn[q], *qk = o->qn[r], *ql = o->qn[s];
    if (qi[0] + qj[0] != qk[0] + ql[0] || qi[1] + qj[1] != qk[1] + ql[1] || qi[2] + qj[2] != qk[2] + ql[2])
        return 0.0;
    const double L = o->inf_L;
    const double V_0R = 200, V_0T = 178, V_0S = 91.85;
    const double kappa_R = 1.487, kappa_T = 0.639, kappa_S = 0.465;
    double qSquared = 0.;
    for (int i = 0; i < 3; i++) {
        double ki = 2 * M_PI * qi[i] / L, kj = 2 * M_PI * qj[i] / L;
        double kk = 2 * M_PI * qk[i] / L, kl = 2 * M_PI * ql[i] / L;
        double bra = 0.5 * (ki - kj), ket = 0.5 * (kk - kl);
        double tr = bra - ket;
        qSquared += tr * tr;
    }
    double V_R = V_0R / (L * L * L) * pow(M_PI / kappa_R, 1.5) * exp(-qSquared / (4 * kappa_R));
    double V_T = -V_0T / (L * L * L) * pow(M_PI / kappa_T, 1.5) * exp(-qSquared / (4 * kappa_T));
    double V_S = -V_0S / (L * L * L) * pow(M_PI / kappa_S, 1.5) * exp(-qSquared / (4 * kappa_S));
    double spinEx = spin_exchange(qi[3], qj[3], qk[3], ql[3]);
    double isoSpinEx = spin_exchange(qi[4], qj[4], qk[4], ql[4]);
    double IsIt = kron_del(qi[3], qk[3]) * kron_del(qj[3], ql[3]) * kron_del(qi[4], qk[4]) * kron_del(qj[4], ql[4]);
    double PsIt = spinEx * kron_del(qi[4], qk[4]) * kron_del(qj[4], ql[4]);
    double PsPt = spinEx * isoSpinEx;
    double IsPt = kron_del(qi[3], qk[3]) * kron_del(qj[3], ql[3]) * isoSpinEx;
    return 0.5 * (V_R + 0.5 * V_T + 0.5 * V_S) * IsIt + 0.25 * (V_T - V_S) * PsIt -
           0.5 * (V_R + 0.5 * V_T + 0.5 * V_S) * PsPt - 0.25 * (V_T - V_S) * IsPt;
}

/* calcTBME: antisymmetrised matrix element (ElectronGasSPBasis.cpp:419-422,
 * InfMatterSPBasis.cpp:355-359). */
static double tbme(const orc_t *o, size_t p, size_t q, size_t r, size_t s) {
    if (o->model == 0) return pairing_tbme(o, p, q, r, s);
    if (o->model == 1) return minnesota_direct(o, p, q, r, s) - minnesota_direct(o, p, q, s, r);
    return heg_direct(o, p, q, r, s) - heg_direct(o, p, q, s, r);
}

/* ------------------------------------------------------------- channels */

static size_t raw_chan(const orc_t *o, size_t p, size_t q) {
    if (o->model == 0) /* PairingSPBasis.cpp:168-174 */
        return (size_t)((o->qn[p][1] + o->qn[q][1] + 2) / 2);
    /* ElectronGasSPBasis.cpp:572-596 (one species: TBtMax = 1, tz index 0) */
    size_t tb = 4 * (size_t)o->nmax + 1;
    long nx = o->qn[p][0] + o->qn[q][0] + (long)(tb / 2);
    long ny = o->qn[p][1] + o->qn[q][1] + (long)(tb / 2);
    long nz = o->qn[p][2] + o->qn[q][2] + (long)(tb / 2);
    long sz = (o->qn[p][3] + o->qn[q][3] + 2) / 2;
    return (size_t)nx * tb * tb * 3 + (size_t)ny * tb * 3 + (size_t)nz * 3 + (size_t)sz;
}

static size_t raw_mod(const orc_t *o, size_t p, size_t q) {
    if (o->model == 0) /* PairingSPBasis.cpp:176-181 */
        return (size_t)((o->qn[p][1] - o->qn[q][1] + 2) / 2);
    /* ElectronGasSPBasis.cpp:599-624 */
    size_t tb = 4 * (size_t)o->nmax + 1;
    long nx = o->qn[p][0] - o->qn[q][0] + (long)(tb / 2);
    long ny = o->qn[p][1] - o->qn[q][1] + (long)(tb / 2);
    long nz = o->qn[p][2] - o->qn[q][2] + (long)(tb / 2);
    long sz = (o->qn[p][3] - o->qn[q][3] + 2) / 2;
    return (size_t)nx * tb * tb * 3 + (size_t)ny * tb * 3 + (size_t)nz * 3 + (size_t)sz;
}

typedef struct { size_t index; double cost; } load_t;
static int load_cmp(const void *x, const void *y) {
    const load_t *a = x, *b = y;
    if (a->cost > b->cost) return -1;
    if (a->cost < b->cost) return 1;
    return (a->index < b->index) ? -1 : (a->index > b->index);
}

static int triple_cmp(const void *x, const void *y) {
    const otriple *a = x, *b = y;
    if (a->p != b->p) return a->p < b->p ? -1 : 1;
    if (a->q != b->q) return a->q < b->q ? -1 : 1;
    if (a->r != b->r) return a->r < b->r ? -1 : 1;
    return 0;
}

/* Normal ordering and E_ref (SPBasis.cpp:7-40). */
static void reference_and_normal_order(orc_t *o) {
    double e = 0.0;
    for (size_t i = 0; i < o->nh; i++)
        for (size_t j = 0; j < o->nh; j++) e += 0.5 * tbme(o, i, j, i, j);
    for (size_t i = 0; i < o->nh; i++) e += o->eps[i];
    o->eref = e;
    for (size_t i = 0; i < o->nh; i++) {
        double ei = o->eps[i];
        for (size_t j = 0; j < o->nh; j++) ei += tbme(o, i, j, i, j);
        o->eps[i] = ei;
    }
    for (size_t a = o->nh; a < o->nsp; a++) {
        double ea = o->eps[a];
        for (size_t j = 0; j < o->nh; j++) ea += tbme(o, a, j, a, j);
        o->eps[a] = ea;
    }
}

/* Channel dims, ffd order at one rank, sorted pair maps and 3-body lists
 * (main.cpp:78-131, balance.cpp:91-144, SPBasis.cpp:80-819). */
static void setup_channels(orc_t *o) {
    size_t n = o->nch, nsp = o->nsp, nh = o->nh;
    size_t *pp = calloc(n, sizeof(size_t)), *hh = calloc(n, sizeof(size_t));
    size_t *hp = calloc(n, sizeof(size_t)), *ph = calloc(n, sizeof(size_t));
    size_t *mph = calloc(n, sizeof(size_t)), *mhp = calloc(n, sizeof(size_t));
    for (size_t p = 0; p < nsp; p++)
        for (size_t q = 0; q < nsp; q++) {
            size_t c = raw_chan(o, p, q), m = raw_mod(o, p, q);
            if (p < nh) { if (q < nh) hh[c]++; else { hp[c]++; mhp[m]++; } }
            else { if (q < nh) { ph[c]++; mph[m]++; } else pp[c]++; }
        }
    /* cost model of main.cpp:94-112 (direct dims only) */
    load_t *ld = malloc(n * sizeof(load_t));
    for (size_t c = 0; c < n; c++) {
        double P = (double)pp[c], H = (double)hh[c], PH = (double)ph[c], HP = (double)hp[c];
        double cost = 0;
        cost += H * H; cost += H * HP; cost += H * P; cost += HP * PH;
        cost += HP * P; cost += P * P; cost += HP * PH; cost += HP * HP;
        ld[c].index = c; ld[c].cost = cost;
    }
    qsort(ld, n, sizeof(load_t), load_cmp);
    o->ch_raw = malloc(n * sizeof(size_t));
    o->ch_inv = malloc(n * sizeof(size_t));
    for (size_t c = 0; c < n; c++) { o->ch_raw[c] = ld[c].index; o->ch_inv[ld[c].index] = c; }
    free(ld);
    o->pp = malloc(n * sizeof(size_t)); o->hh = malloc(n * sizeof(size_t));
    o->mph = malloc(n * sizeof(size_t)); o->mhp = malloc(n * sizeof(size_t));
    for (size_t c = 0; c < n; c++) {
        size_t r = o->ch_raw[c];
        o->pp[c] = pp[r]; o->hh[c] = hh[r]; o->mph[c] = mph[r]; o->mhp[c] = mhp[r];
    }
    free(pp); free(hh); free(hp); free(ph); free(mph); free(mhp);

    o->ppl = calloc(n, sizeof(opair *)); o->hhl = calloc(n, sizeof(opair *));
    o->mphl = calloc(n, sizeof(opair *)); o->mhpl = calloc(n, sizeof(opair *));
    size_t *fp = calloc(n, sizeof(size_t)), *fh = calloc(n, sizeof(size_t));
    size_t *fmp = calloc(n, sizeof(size_t)), *fmh = calloc(n, sizeof(size_t));
    for (size_t c = 0; c < n; c++) {
        o->ppl[c] = malloc((o->pp[c] + 1) * sizeof(opair));
        o->hhl[c] = malloc((o->hh[c] + 1) * sizeof(opair));
        o->mphl[c] = malloc((o->mph[c] + 1) * sizeof(opair));
        o->mhpl[c] = malloc((o->mhp[c] + 1) * sizeof(opair));
    }
    /* lexicographic enumeration gives the std::sort order of SPBasis.cpp:372-378 */
    for (size_t p = 0; p < nsp; p++)
        for (size_t q = 0; q < nsp; q++) {
            size_t c = o->ch_inv[raw_chan(o, p, q)], m = o->ch_inv[raw_mod(o, p, q)];
            opair pq = {(int32_t)p, (int32_t)q};
            if (p < nh && q < nh) o->hhl[c][fh[c]++] = pq;
            if (p >= nh && q >= nh) o->ppl[c][fp[c]++] = pq;
            if (p >= nh && q < nh) o->mphl[m][fmp[m]++] = pq;
            if (p < nh && q >= nh) o->mhpl[m][fmh[m]++] = pq;
        }
    free(fp); free(fh); free(fmp); free(fmh);

    /* 3-body lists hpp(x) = {(j,a,b)} and phh(a) = {(b,i,j)} (SPBasis.cpp:180-244,
     * 463-565): the T2 elements with first hole x / first particle a. */
    o->hpp_dim = calloc(nh, sizeof(size_t));
    o->phh_dim = calloc(nsp - nh, sizeof(size_t));
    for (size_t c = 0; c < n; c++)
        for (size_t k = 0; k < o->hh[c]; k++) o->hpp_dim[o->hhl[c][k].p] += o->pp[c];
    for (size_t c = 0; c < n; c++)
        for (size_t k = 0; k < o->pp[c]; k++) o->phh_dim[o->ppl[c][k].p - nh] += o->hh[c];
    o->hppl = malloc(nh * sizeof(otriple *));
    o->phhl = malloc((nsp - nh) * sizeof(otriple *));
    size_t *fill = calloc(nsp, sizeof(size_t));
    for (size_t x = 0; x < nh; x++) o->hppl[x] = malloc((o->hpp_dim[x] + 1) * sizeof(otriple));
    for (size_t a = 0; a < nsp - nh; a++) o->phhl[a] = malloc((o->phh_dim[a] + 1) * sizeof(otriple));
    for (size_t c = 0; c < n; c++)
        for (size_t k = 0; k < o->hh[c]; k++)
            for (size_t l = 0; l < o->pp[c]; l++) {
                int32_t x = o->hhl[c][k].p, j = o->hhl[c][k].q;
                int32_t a = o->ppl[c][l].p, b = o->ppl[c][l].q;
                otriple t1 = {j, a, b};
                o->hppl[x][fill[x]++] = t1;
                otriple t2 = {b, x, j};
                o->phhl[a - nh][fill[a]++] = t2;
            }
    free(fill);
    #pragma omp parallel for schedule(dynamic)
    for (size_t x = 0; x < nsp; x++) {
        if (x < nh) qsort(o->hppl[x], o->hpp_dim[x], sizeof(otriple), triple_cmp);
        else qsort(o->phhl[x - nh], o->phh_dim[x - nh], sizeof(otriple), triple_cmp);
    }

    /* chain offsets (Chain.hpp:405-409) */
    o->off_t = malloc((n + 1) * sizeof(size_t));
    o->off_hhhh = malloc((n + 1) * sizeof(size_t));
    o->off_pppp = malloc((n + 1) * sizeof(size_t));
    o->off_ph = malloc((n + 1) * sizeof(size_t));
    o->off_mb = malloc((n + 1) * sizeof(size_t));
    o->off_t[0] = o->off_hhhh[0] = o->off_pppp[0] = o->off_ph[0] = o->off_mb[0] = 0;
    for (size_t c = 0; c < n; c++) {
        o->off_t[c + 1] = o->off_t[c] + o->pp[c] * o->hh[c];
        o->off_hhhh[c + 1] = o->off_hhhh[c] + o->hh[c] * o->hh[c];
        o->off_pppp[c + 1] = o->off_pppp[c] + o->pp[c] * o->pp[c];
        o->off_ph[c + 1] = o->off_ph[c] + o->mph[c] * o->mhp[c];
        o->off_mb[c + 1] = o->off_mb[c] + o->mhp[c] * o->mhp[c];
    }
    o->nt2 = o->off_t[n]; o->nhhhh = o->off_hhhh[n]; o->npppp = o->off_pppp[n];
    o->nmb = o->off_mb[n];
    o->off_hpp = malloc((nh + 1) * sizeof(size_t));
    o->off_phh = malloc((nsp - nh + 1) * sizeof(size_t));
    o->off_hpp[0] = 0; o->off_phh[0] = 0;
    for (size_t x = 0; x < nh; x++) o->off_hpp[x + 1] = o->off_hpp[x] + o->hpp_dim[x];
    for (size_t a = 0; a < nsp - nh; a++) o->off_phh[a + 1] = o->off_phh[a] + o->phh_dim[a];
}

/* ------------------------------------------------------------- bases */

orc_t *orc_create_pairing(int nPairStates, int nParticles, double g, double xi) {
    orc_t *o = calloc(1, sizeof(orc_t));
    o->model = 0;
    o->g = g; o->xi = xi;
    o->nsp = 2 * (size_t)nPairStates;
    o->nh = (size_t)nParticles;
    o->nch = 3;
    o->qn = calloc(o->nsp, sizeof(*o->qn));
    o->eps = malloc(o->nsp * sizeof(double));
    /* PairingSPBasis.cpp:27-58: state 2l = (l+1, +1), 2l+1 = (l+1, -1) */
    for (size_t l = 0; l < (size_t)nPairStates; l++) {
        o->qn[2 * l][0] = (int)l + 1; o->qn[2 * l][1] = +1;
        o->qn[2 * l + 1][0] = (int)l + 1; o->qn[2 * l + 1][1] = -1;
    }
    for (size_t i = 0; i < o->nsp; i++) o->eps[i] = (o->qn[i][0] - 1.0) * xi;
    reference_and_normal_order(o);
    setup_channels(o);
    return o;
}

/* Plane-wave momenta n^2 <= EMax with two spin states and one isospin state
 * (tz = -1): the index maps of ElectronGasSPBasis.cpp:38-170 and, for
 * pure neutron matter, InfMatterSPBasis.cpp:14-163 (same loops and order). */
static void plane_wave_basis(orc_t *o, int nShells, int nParticleShells);

orc_t *orc_create_heg(double r_s, int nShells, int nParticleShells) {
    orc_t *o = calloc(1, sizeof(orc_t));
    o->model = 2;
    o->r_s = r_s;
    plane_wave_basis(o, nShells, nParticleShells);
    const size_t nsp = o->nsp, nh = o->nh;
    const int nmax = o->nmax;
    /* generateBasis (ElectronGasSPBasis.cpp:173-210): reference constants */
    double hbarc = 0.1973269788;
    double eVs_in_a_Hartree = 27.21138505;
    hbarc = hbarc * 10000 / eVs_in_a_Hartree;
    double massc2 = 0.5109989461;
    massc2 = massc2 * 1000000 / eVs_in_a_Hartree;
    double fine_struc = 1.0 / 137.035999139;
    double r_bohr = hbarc / (massc2 * fine_struc);
    double r_0 = r_s * r_bohr;
    double prefactor = hbarc * hbarc / (2. * massc2);
    o->L = pow((4. / 3.) * M_PI * (double)nh, 1. / 3.) * r_0;
    o->e_sq = fine_struc * hbarc;
    o->eps = malloc(nsp * sizeof(double));
    for (size_t i = 0; i < nsp; i++) {
        int *q = o->qn[i];
        o->eps[i] = (prefactor * 4 * M_PI * M_PI / (o->L * o->L)) *
                    (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    }
    /* ElectronGasSPBasis.cpp:256-262: (4 nMax + 1)^3 x 3 channels */
    size_t tb = 4 * (size_t)nmax + 1;
    o->nch = tb * tb * tb * 3;
    reference_and_normal_order(o);
    setup_channels(o);
    return o;
}

/* Pure neutron matter with the Minnesota potential
 * (InfMatterSPBasis(1, density, tzMax = 1, nShells, nParticleShells)). */
orc_t *orc_create_infmatter(double density, int nShells, int nParticleShells) {
    orc_t *o = calloc(1, sizeof(orc_t));
    o->model = 1;
    o->density = density;
    plane_wave_basis(o, nShells, nParticleShells);
    const size_t nsp = o->nsp;
    /* generateBasis (InfMatterSPBasis.cpp:166-184): neutron mass, L from the density */
    const double hbarc = 197.3269788, massc2 = 939.565;
    double prefactor = hbarc * hbarc / (2. * massc2);
    o->L = pow(o->nh / density, 1. / 3.);
    o->inf_L = pow(o->nh / o->density, 1. / 3.);   /* as recomputed in calc_TBME_not_antisym */
    o->eps = malloc(nsp * sizeof(double));
    for (size_t i = 0; i < nsp; i++) {
        int *q = o->qn[i];
        o->eps[i] = (prefactor * 4 * M_PI * M_PI / (o->L * o->L)) *
                    (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    }
    /* setUpChannelValues (InfMatterSPBasis.cpp:196-240) with chanTzMax = -2:
     * (4 nMax + 1)^3 x 3 x 1 channels, keyed like the electron gas */
    size_t tb = 4 * (size_t)o->nmax + 1;
    o->nch = tb * tb * tb * 3;
    reference_and_normal_order(o);
    setup_channels(o);
    return o;
}

static void plane_wave_basis(orc_t *o, int nShells, int nParticleShells) {
    /* ElectronGasSPBasis.cpp:38-96 */
    size_t emax = 0;
    for (size_t i = 1; i < (size_t)nShells; emax++)
        if (three_squares((unsigned)emax + 1)) i++;
    int nmax = (int)sqrt_ceil((unsigned)emax);
    size_t nsp = 0;
    int nmaxa = 0;
    for (int nx = -nmax; nx <= nmax; nx++)
        for (int ny = -nmax; ny <= nmax; ny++)
            for (int nz = -nmax; nz <= nmax; nz++)
                if ((size_t)(nx * nx + ny * ny + nz * nz) <= emax) {
                    if (nx > nmaxa) nmaxa = nx;
                    nsp += 2;
                }
    nmax = nmaxa;
    size_t ef = 0;
    for (size_t i = 1; i < (size_t)nParticleShells; ef++)
        if (three_squares((unsigned)ef + 1)) i++;
    int npmax = (int)sqrt_ceil((unsigned)ef);
    size_t nh = 0;
    for (int nx = -npmax; nx <= npmax; nx++)
        for (int ny = -npmax; ny <= npmax; ny++)
            for (int nz = -npmax; nz <= npmax; nz++)
                if ((size_t)(nx * nx + ny * ny + nz * nz) <= ef) nh += 2;
    o->emax = emax; o->efermi = ef; o->nmax = nmax;
    o->nsp = nsp; o->nh = nh;
    /* generateIndexMap (ElectronGasSPBasis.cpp:109-170): order by n^2, then
     * nx, ny, nz loops, then sz = -1, +1 */
    size_t *shells = calloc(emax + 2, sizeof(size_t));
    for (int nx = -nmax; nx <= nmax; nx++)
        for (int ny = -nmax; ny <= nmax; ny++)
            for (int nz = -nmax; nz <= nmax; nz++) {
                size_t E = (size_t)(nx * nx + ny * ny + nz * nz);
                if (E <= emax) shells[E + 1] += 2;
            }
    for (size_t E = 1; E <= emax; E++) shells[E] += shells[E - 1];
    o->qn = calloc(nsp, sizeof(*o->qn));
    for (int nx = -nmax; nx <= nmax; nx++)
        for (int ny = -nmax; ny <= nmax; ny++)
            for (int nz = -nmax; nz <= nmax; nz++) {
                size_t E = (size_t)(nx * nx + ny * ny + nz * nz);
                if (E > emax) continue;
                for (int sz = -1; sz <= 1; sz += 2) {
                    int *q = o->qn[shells[E]++];
                    q[0] = nx; q[1] = ny; q[2] = nz; q[3] = sz; q[4] = -1;
                }
            }
    free(shells);
}

/* ------------------------------------------------------------- load */

/* loadFreeVnn_pppp (load.cpp:112-318 conventions): one direct channel's
 * V^pppp block, [(a,b),(c,d)] = <ab|v|cd>, row-major. */
static void channel_vpppp(const orc_t *o, size_t c, double *vp) {
    size_t P = o->pp[c];
    for (size_t k = 0; k < P; k++)
        for (size_t l = 0; l < P; l++)
            vp[k * P + l] = tbme(o, o->ppl[c][k].p, o->ppl[c][k].q, o->ppl[c][l].p, o->ppl[c][l].q);
}

void orc_set_matrix_free(orc_t *o, int on) { if (!o->loaded) o->free_pppp = on != 0; }

static size_t find_triple(const otriple *l, size_t n, otriple key) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        int c = triple_cmp(&l[mid], &key);
        if (c == 0) return mid;
        if (c < 0) lo = mid + 1; else hi = mid;
    }
    fprintf(stderr, "oracle: triple not found\n");
    abort();
}

int orc_load(orc_t *o) {
    if (o->loaded) return 0;
    size_t n = o->nch, nsp = o->nsp, nh = o->nh;
    o->v_hhhh = malloc((o->nhhhh + 1) * sizeof(double));
    o->v_hhpp = malloc((o->nt2 + 1) * sizeof(double));
    o->v_pppp = malloc(((o->free_pppp ? 0 : o->npppp) + 1) * sizeof(double));
    o->v_phh_p = malloc((o->nt2 + 1) * sizeof(double));
    o->v_h_hpp = malloc((o->nt2 + 1) * sizeof(double));
    o->v_modA = malloc((o->off_ph[n] + 1) * sizeof(double));
    o->v_modB = malloc((o->nmb + 1) * sizeof(double));
    if (!o->v_pppp || !o->v_hhpp) return -1;
    /* loadFreeVnn_* (load.cpp:112-318) */
    #pragma omp parallel for schedule(dynamic)
    for (size_t c = 0; c < n; c++) {
        size_t P = o->pp[c], H = o->hh[c];
        double *vh = o->v_hhhh + o->off_hhhh[c];
        for (size_t k = 0; k < H; k++)
            for (size_t l = 0; l < H; l++)
                vh[k * H + l] = tbme(o, o->hhl[c][k].p, o->hhl[c][k].q, o->hhl[c][l].p, o->hhl[c][l].q);
        double *vhp = o->v_hhpp + o->off_t[c];
        for (size_t k = 0; k < H; k++)
            for (size_t l = 0; l < P; l++)
                vhp[k * P + l] = tbme(o, o->hhl[c][k].p, o->hhl[c][k].q, o->ppl[c][l].p, o->ppl[c][l].q);
        if (!o->free_pppp) channel_vpppp(o, c, o->v_pppp + o->off_pppp[c]);
        /* mod channel c: V^A hp x ph = <ij|v|ab> for (i,a),(b,j) (load.cpp:275-296),
         * V^B hp x hp = <ia|v|bj> for (i,b),(j,a) (load.cpp:298-318) */
        size_t MP = o->mph[c], MH = o->mhp[c];
        double *va = o->v_modA + o->off_ph[c];
        for (size_t k = 0; k < MH; k++)
            for (size_t l = 0; l < MP; l++) {
                int32_t i = o->mhpl[c][k].p, a = o->mhpl[c][k].q;
                int32_t b = o->mphl[c][l].p, j = o->mphl[c][l].q;
                va[k * MP + l] = tbme(o, i, j, a, b);
            }
        double *vb = o->v_modB + o->off_mb[c];
        for (size_t k = 0; k < MH; k++)
            for (size_t l = 0; l < MH; l++) {
                int32_t i = o->mhpl[c][k].p, b = o->mhpl[c][k].q;
                int32_t j = o->mhpl[c][l].p, a = o->mhpl[c][l].q;
                vb[k * MH + l] = tbme(o, i, a, b, j);
            }
    }
    #pragma omp parallel for schedule(dynamic)
    for (size_t x = 0; x < nsp; x++) {
        if (x < nh) { /* loadFreeVnn_h_hpp (load.cpp:258-273) */
            double *v = o->v_h_hpp + o->off_hpp[x];
            for (size_t k = 0; k < o->hpp_dim[x]; k++) {
                otriple t = o->hppl[x][k];
                v[k] = tbme(o, x, t.p, t.q, t.r);
            }
        } else { /* loadFreeVnn_phh_p (load.cpp:241-256) */
            size_t a = x - nh;
            double *v = o->v_phh_p + o->off_phh[a];
            for (size_t k = 0; k < o->phh_dim[a]; k++) {
                otriple t = o->phhl[a][k];
                v[k] = tbme(o, t.q, t.r, x, t.p);
            }
        }
    }
    /* loadMaps (load.cpp:389-532) */
    size_t *pos_mod = malloc(nsp * nsp * sizeof(size_t));
    for (size_t m = 0; m < n; m++) {
        for (size_t k = 0; k < o->mph[m]; k++) pos_mod[o->mphl[m][k].p * nsp + o->mphl[m][k].q] = k;
        for (size_t k = 0; k < o->mhp[m]; k++) pos_mod[o->mhpl[m][k].p * nsp + o->mhpl[m][k].q] = k;
    }
    size_t nt = o->nt2;
    o->m_iab_j = malloc((nt + 1) * sizeof(size_t)); o->m_jab_i = malloc((nt + 1) * sizeof(size_t));
    o->m_a_bij = malloc((nt + 1) * sizeof(size_t)); o->m_b_aij = malloc((nt + 1) * sizeof(size_t));
    o->m_ai_jb = malloc((nt + 1) * sizeof(size_t)); o->m_aj_ib = malloc((nt + 1) * sizeof(size_t));
    o->m_bi_ja = malloc((nt + 1) * sizeof(size_t)); o->m_bj_ia = malloc((nt + 1) * sizeof(size_t));
    #pragma omp parallel for schedule(dynamic)
    for (size_t c = 0; c < n; c++) {
        size_t P = o->pp[c], H = o->hh[c];
        for (size_t AB = 0; AB < P; AB++)
            for (size_t IJ = 0; IJ < H; IJ++) {
                size_t e = o->off_t[c] + AB * H + IJ;
                int32_t a = o->ppl[c][AB].p, b = o->ppl[c][AB].q;
                int32_t i = o->hhl[c][IJ].p, j = o->hhl[c][IJ].q;
                otriple k1 = {i, a, b}, k2 = {j, a, b}, k3 = {b, i, j}, k4 = {a, i, j};
                o->m_iab_j[e] = o->off_hpp[j] + find_triple(o->hppl[j], o->hpp_dim[j], k1);
                o->m_jab_i[e] = o->off_hpp[i] + find_triple(o->hppl[i], o->hpp_dim[i], k2);
                o->m_a_bij[e] = o->off_phh[a - nh] + find_triple(o->phhl[a - nh], o->phh_dim[a - nh], k3);
                o->m_b_aij[e] = o->off_phh[b - nh] + find_triple(o->phhl[b - nh], o->phh_dim[b - nh], k4);
                size_t m;
                m = o->ch_inv[raw_mod(o, a, i)];
                o->m_ai_jb[e] = o->off_ph[m] + pos_mod[a * nsp + i] * o->mhp[m] + pos_mod[j * nsp + b];
                m = o->ch_inv[raw_mod(o, a, j)];
                o->m_aj_ib[e] = o->off_ph[m] + pos_mod[a * nsp + j] * o->mhp[m] + pos_mod[i * nsp + b];
                m = o->ch_inv[raw_mod(o, b, i)];
                o->m_bi_ja[e] = o->off_ph[m] + pos_mod[b * nsp + i] * o->mhp[m] + pos_mod[j * nsp + a];
                m = o->ch_inv[raw_mod(o, b, j)];
                o->m_bj_ia[e] = o->off_ph[m] + pos_mod[b * nsp + j] * o->mhp[m] + pos_mod[i * nsp + a];
            }
    }
    free(pos_mod);
    /* work chains (main.cpp:238-264) */
    size_t nb = (nt + 1) * sizeof(double);
    o->t2 = calloc(1, nb); o->t2old = calloc(1, nb); o->d = calloc(1, nb);
    o->r = calloc(1, nb); o->tph = calloc(1, nb);
    o->s_jab_i = calloc(1, nb); o->s_iab_j = calloc(1, nb);
    o->s_a_bij = calloc(1, nb); o->s_b_aij = calloc(1, nb);
    o->diffh = calloc(1, nb); o->diffp = calloc(1, nb);
    for (int k = 0; k < 8; k++) o->gat[k] = calloc(1, nb);
    o->chan_e = calloc(n + 1, sizeof(double));
    o->loaded = 1;
    return 0;
}

/* ------------------------------------------------------------- accessors */

size_t orc_nsp(const orc_t *o) { return o->nsp; }
size_t orc_nholes(const orc_t *o) { return o->nh; }
size_t orc_nchannels(const orc_t *o) { return o->nch; }
size_t orc_nt2(const orc_t *o) { return o->nt2; }
double orc_eref(const orc_t *o) { return o->eref; }
void orc_get_eps(const orc_t *o, double *eps) { memcpy(eps, o->eps, o->nsp * sizeof(double)); }
void orc_get_qn(const orc_t *o, int32_t *qn) {
    for (size_t p = 0; p < o->nsp; p++)
        for (int k = 0; k < 5; k++) qn[5 * p + k] = o->qn[p][k];
}

void orc_get_dims(const orc_t *o, int64_t *pp, int64_t *hh, int64_t *mph, int64_t *mhp) {
    for (size_t c = 0; c < o->nch; c++) {
        pp[c] = (int64_t)o->pp[c]; hh[c] = (int64_t)o->hh[c];
        mph[c] = (int64_t)o->mph[c]; mhp[c] = (int64_t)o->mhp[c];
    }
}

void orc_t2_labels(const orc_t *o, int32_t *a, int32_t *b, int32_t *i, int32_t *j) {
    for (size_t c = 0; c < o->nch; c++)
        for (size_t AB = 0; AB < o->pp[c]; AB++)
            for (size_t IJ = 0; IJ < o->hh[c]; IJ++) {
                size_t e = o->off_t[c] + AB * o->hh[c] + IJ;
                a[e] = o->ppl[c][AB].p; b[e] = o->ppl[c][AB].q;
                i[e] = o->hhl[c][IJ].p; j[e] = o->hhl[c][IJ].q;
            }
}

void orc_tph_labels(const orc_t *o, int32_t *a, int32_t *i, int32_t *j, int32_t *b) {
    for (size_t m = 0; m < o->nch; m++)
        for (size_t r = 0; r < o->mph[m]; r++)
            for (size_t c = 0; c < o->mhp[m]; c++) {
                size_t e = o->off_ph[m] + r * o->mhp[m] + c;
                a[e] = o->mphl[m][r].p; i[e] = o->mphl[m][r].q;
                j[e] = o->mhpl[m][c].p; b[e] = o->mhpl[m][c].q;
            }
}

void orc_get_vhhpp(const orc_t *o, double *v) { memcpy(v, o->v_hhpp, o->nt2 * sizeof(double)); }

/* ------------------------------------------------------------- diagrams */

double orc_mbpt2(const orc_t *o) {
    /* MBPT2ChannelCorr / MBPT2Corr (MBPT2Corr.cpp:6-38): serial channel loop */
    double corr = 0.0;
    for (size_t c = 0; c < o->nch; c++) {
        size_t P = o->pp[c], H = o->hh[c];
        const double *v = o->v_hhpp + o->off_t[c];
        double cc = 0.0;
        for (size_t k = 0; k < H; k++)
            for (size_t l = 0; l < P; l++) {
                size_t h0 = o->hhl[c][k].p, h1 = o->hhl[c][k].q;
                size_t p0 = o->ppl[c][l].p, p1 = o->ppl[c][l].q;
                double den = o->eps[h0] + o->eps[h1] - o->eps[p0] - o->eps[p1];
                double x = v[k * P + l];
                cc += x * x / den;
            }
        corr += 0.25 * cc;
    }
    return corr;
}

void orc_ladder(const orc_t *o, const double *t, double *d) {
    /* calcT2Diff_pphh (calc.cpp:132-231); with free_pppp the per-thread V^pppp
     * buffer of calcFreeT2Diff_pphh (calc.cpp:233-262) */
    #pragma omp parallel
    {
        size_t mx = 1, mp = 1;
        for (size_t c = 0; c < o->nch; c++) {
            if (o->hh[c] * o->hh[c] > mx) mx = o->hh[c] * o->hh[c];
            if (o->free_pppp && o->pp[c] * o->pp[c] > mp) mp = o->pp[c] * o->pp[c];
        }
        double *x = malloc(mx * sizeof(double));
        double *vbuf = o->free_pppp ? malloc(mp * sizeof(double)) : NULL;
        #pragma omp for schedule(dynamic)
        for (size_t c = 0; c < o->nch; c++) {
            size_t P = o->pp[c], H = o->hh[c];
            if (!P || !H) continue;
            memcpy(x, o->v_hhhh + o->off_hhhh[c], H * H * sizeof(double));
            const double *tc = t + o->off_t[c];
            double *dc = d + o->off_t[c];
            dgemm_rm(H, H, P, 0.5, o->v_hhpp + o->off_t[c], tc, 1.0, x);
            dgemm_rm(P, H, H, 0.5, tc, x, 0.0, dc);
            const double *vp = o->v_pppp + o->off_pppp[c];
            if (vbuf) { channel_vpppp(o, c, vbuf); vp = vbuf; }
            dgemm_rm(P, H, P, 0.5, vp, tc, 1.0, dc);
        }
        free(x);
        free(vbuf);
    }
}

void orc_ring(const orc_t *o, const double *tph, double *r) {
    /* calcT2Diff_phhp (calc.cpp:264-339) */
    #pragma omp parallel
    {
        size_t mx = 1;
        for (size_t c = 0; c < o->nch; c++) if (o->mhp[c] * o->mhp[c] > mx) mx = o->mhp[c] * o->mhp[c];
        double *x = malloc(mx * sizeof(double));
        #pragma omp for schedule(dynamic)
        for (size_t m = 0; m < o->nch; m++) {
            size_t PH = o->mph[m], HP = o->mhp[m];
            if (!PH || !HP) continue;
            memcpy(x, o->v_modB + o->off_mb[m], HP * HP * sizeof(double));
            const double *tm = tph + o->off_ph[m];
            dgemm_rm(HP, HP, PH, 0.5, o->v_modA + o->off_ph[m], tm, 1.0, x);
            dgemm_rm(PH, HP, HP, 1.0, tm, x, 0.0, r + o->off_ph[m]);
        }
        free(x);
    }
}

void orc_scatter_ph(const orc_t *o, const double *t, double *tph) {
    for (size_t e = 0; e < o->nt2; e++) tph[o->m_ai_jb[e]] = t[e];
}

/* Scatter T (pphh) into the 3-body layouts and T^ph (main.cpp:661-665). */
static void scatter_all(orc_t *o, const double *t) {
    #pragma omp parallel for schedule(static)
    for (size_t e = 0; e < o->nt2; e++) {
        o->s_jab_i[o->m_jab_i[e]] = t[e];
        o->s_iab_j[o->m_iab_j[e]] = t[e];
        o->s_a_bij[o->m_a_bij[e]] = t[e];
        o->s_b_aij[o->m_b_aij[e]] = t[e];
        o->tph[o->m_ai_jb[e]] = t[e];
    }
}

/* calcChannelT2Diff_h2pp_h1 / _p2_p1hh (calc.cpp:9-107), keeping the argument
 * order of the main.cpp:591-606 calls: chi uses t2Old_h2pp_h1 (= S_jab_i),
 * the hole term multiplies t2Old_h1pp_h2 (= S_iab_j); psi uses
 * t2Old_p1_p2hh (= S_a_bij), the particle term multiplies t2Old_p2_p1hh
 * (= S_b_aij). */
static void onebody(orc_t *o, double *chi, double *psi) {
    size_t nh = o->nh, np = o->nsp - o->nh;
    #pragma omp parallel for schedule(dynamic)
    for (size_t a = 0; a < np; a++) {
        const double *v = o->v_phh_p + o->off_phh[a];
        const double *s1 = o->s_a_bij + o->off_phh[a];
        const double *s2 = o->s_b_aij + o->off_phh[a];
        double value = 0.0;
        for (size_t k = 0; k < o->phh_dim[a]; k++) value += s1[k] * v[k];
        double *dp = o->diffp + o->off_phh[a];
        for (size_t k = 0; k < o->phh_dim[a]; k++) dp[k] = -0.5 * value * s2[k];
        if (psi) psi[a] = value;
    }
    #pragma omp parallel for schedule(dynamic)
    for (size_t i = 0; i < nh; i++) {
        const double *v = o->v_h_hpp + o->off_hpp[i];
        const double *s1 = o->s_jab_i + o->off_hpp[i];
        const double *s2 = o->s_iab_j + o->off_hpp[i];
        double value = 0.0;
        for (size_t k = 0; k < o->hpp_dim[i]; k++) value += v[k] * s1[k];
        double *dh = o->diffh + o->off_hpp[i];
        for (size_t k = 0; k < o->hpp_dim[i]; k++) dh[k] = 0.5 * value * s2[k];
        if (chi) chi[i] = value;
    }
}

void orc_onebody(const orc_t *o_, const double *t, double *chi, double *psi) {
    orc_t *o = (orc_t *)o_;
    scatter_all(o, t);
    onebody(o, chi, psi);
}

/* One iteration body (main.cpp:589-670) on o->t2old and the scattered
 * layouts; writes o->t2 and returns the energy of the new amplitudes. */
static double iterate(orc_t *o, double mixing) {
    size_t n = o->nch;
    onebody(o, NULL, NULL);                    /* calc.cpp:27-43, 91-107 */
    orc_ring(o, o->tph, o->r);                 /* calc.cpp:315-339 */
    orc_ladder(o, o->t2old, o->d);             /* calc.cpp:205-231 */
    #pragma omp parallel for schedule(static)  /* 8 x mapFromDistributed, main.cpp:632-639 */
    for (size_t e = 0; e < o->nt2; e++) {
        o->gat[0][e] = o->diffh[o->m_iab_j[e]];
        o->gat[1][e] = o->diffh[o->m_jab_i[e]];
        o->gat[2][e] = o->diffp[o->m_a_bij[e]];
        o->gat[3][e] = o->diffp[o->m_b_aij[e]];
        o->gat[4][e] = o->r[o->m_ai_jb[e]];
        o->gat[5][e] = o->r[o->m_aj_ib[e]];
        o->gat[6][e] = o->r[o->m_bi_ja[e]];
        o->gat[7][e] = o->r[o->m_bj_ia[e]];
    }
    /* calcT2_pphh (calc.cpp:368-459) */
    #pragma omp parallel for schedule(dynamic)
    for (size_t c = 0; c < n; c++) {
        size_t P = o->pp[c], H = o->hh[c];
        const double *v = o->v_hhpp + o->off_t[c];
        double corr = 0.0;
        for (size_t AB = 0; AB < P; AB++)
            for (size_t IJ = 0; IJ < H; IJ++) {
                size_t a = o->ppl[c][AB].p, b = o->ppl[c][AB].q;
                size_t i = o->hhl[c][IJ].p, j = o->hhl[c][IJ].q;
                size_t e = o->off_t[c] + AB * H + IJ;
                double value = v[IJ * P + AB];
                value += o->d[e];
                value += o->gat[4][e];
                value -= o->gat[5][e];
                value -= o->gat[6][e];
                value += o->gat[7][e];
                value += o->gat[1][e];
                value -= o->gat[0][e];
                value += o->gat[3][e];
                value -= o->gat[2][e];
                value = value / (o->eps[i] + o->eps[j] - o->eps[a] - o->eps[b]);
                value = mixing * value + (1. - mixing) * o->t2old[e];
                o->t2[e] = value;
                corr += v[IJ * P + AB] * o->t2[e];
            }
        o->chan_e[c] = 0.25 * corr;
    }
    double corr = 0.0;
    for (size_t c = 0; c < n; c++) corr += o->chan_e[c];
    memcpy(o->t2old, o->t2, o->nt2 * sizeof(double));   /* t2Old_pphh.set */
    scatter_all(o, o->t2);                               /* 5 x mapToDistributed */
    return corr;
}

double orc_step(orc_t *o, const double *t_old, double *t_new, double mixing) {
    memcpy(o->t2old, t_old, o->nt2 * sizeof(double));
    scatter_all(o, o->t2old);
    double e = iterate(o, mixing);
    memcpy(t_new, o->t2, o->nt2 * sizeof(double));
    return e;
}

int orc_solve(orc_t *o, double tol, double *energies, int cap, int *stop, double *seconds) {
    /* main.cpp:491-682 */
    double energyDiff = 1.0, corrCCD = 100.0, prev = 100.0;
    unsigned long it = 0;
    int max_iters = 100;
    double mixing = 0.5;
    memset(o->t2, 0, o->nt2 * sizeof(double));
    memset(o->t2old, 0, o->nt2 * sizeof(double));
    scatter_all(o, o->t2old);
    *stop = 0;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    while (fabs(energyDiff) > tol) {
        if (it > (unsigned long)max_iters) { *stop = 1; break; }
        corrCCD = iterate(o, mixing);
        energyDiff = corrCCD - prev;
        prev = corrCCD;
        if ((int)it < cap) energies[it] = corrCCD;
        it++;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (*stop == 0 && isnan(energyDiff)) *stop = 2;
    if (seconds) *seconds = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    return (int)it;
}

void orc_get_t2(const orc_t *o, double *t) { memcpy(t, o->t2, o->nt2 * sizeof(double)); }

/* ------------------------------------------------------------- destroy */

void orc_destroy(orc_t *o) {
    if (!o) return;
    for (size_t c = 0; c < o->nch; c++) {
        free(o->ppl[c]); free(o->hhl[c]); free(o->mphl[c]); free(o->mhpl[c]);
    }
    for (size_t x = 0; x < o->nh; x++) free(o->hppl[x]);
    for (size_t a = 0; a < o->nsp - o->nh; a++) free(o->phhl[a]);
    free(o->ppl); free(o->hhl); free(o->mphl); free(o->mhpl); free(o->hppl); free(o->phhl);
    free(o->hpp_dim); free(o->phh_dim); free(o->ch_raw); free(o->ch_inv);
    free(o->pp); free(o->hh); free(o->mph); free(o->mhp);
    free(o->off_t); free(o->off_hhhh); free(o->off_pppp); free(o->off_ph); free(o->off_mb);
    free(o->off_hpp); free(o->off_phh);
    free(o->v_hhhh); free(o->v_hhpp); free(o->v_pppp); free(o->v_phh_p); free(o->v_h_hpp);
    free(o->v_modA); free(o->v_modB);
    free(o->m_iab_j); free(o->m_jab_i); free(o->m_a_bij); free(o->m_b_aij);
    free(o->m_ai_jb); free(o->m_aj_ib); free(o->m_bi_ja); free(o->m_bj_ia);
    free(o->t2); free(o->t2old); free(o->d); free(o->r); free(o->tph);
    free(o->s_jab_i); free(o->s_iab_j); free(o->s_a_bij); free(o->s_b_aij);
    free(o->diffh); free(o->diffp);
    for (int k = 0; k < 8; k++) free(o->gat[k]);
    free(o->chan_e);
    free(o->qn); free(o->eps);
    free(o);
}
