/*
 * ccd_oracle.h -- CPU restatement of the reference CCD amplitude iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * MI355X path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in coupled_cluster_doubles_amd/
 * links or calls it.
 *
 * It restates, in plain C with OpenMP over channels, the algorithm of the
 * reference code justinlietz/coupled_cluster_doubles:
 *   src/PairingSPBasis.cpp, src/ElectronGasSPBasis.cpp  (bases, TBMEs)
 *   src/SPBasis.cpp                                     (E_ref, normal ordering,
 *                                                        channel dims and maps)
 *   src/balance.cpp:91-144 + src/main.cpp:94-115         (channel order, 1 rank)
 *   src/load.cpp                                        (V blocks, T2 maps)
 *   src/calc.cpp, src/MBPT2Corr.cpp                      (diagrams, energies)
 *   src/main.cpp:491-682                                (iteration protocol)
 * Storage follows the reference Chain<T> layout (src/Chain.hpp:373-426):
 * one contiguous array of row-major blocks, channels in the permuted order.
 *
 * How the oracle is pinned (tests/test_oracle_*.py):
 *   - pairing closed forms for E_ref and MBPT2;
 *   - an independent dense spin-orbital CCD in numpy (oracle/dense_ccd.py) on
 *     the pairing model and on tiny electron-gas bases, iteration by
 *     iteration;
 *   - channel-histogram counts of the electron-gas configs;
 *   - first iterate == MBPT2/2, antisymmetry of T2, dgemm against numpy.
 */
#ifndef CCD_ORACLE_H
#define CCD_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* Pairing model: nPairStates doubly degenerate levels, nParticles particles
 * (reference PairingSPBasis ctor, src/PairingSPBasis.cpp:14-24). */
orc_t *orc_create_pairing(int nPairStates, int nParticles, double g, double xi);
/* Electron gas: reference ElectronGasSPBasis(2, r_s, 1, nShells,
 * nParticleShells) (src/parse.cpp:239). */
orc_t *orc_create_heg(double r_s, int nShells, int nParticleShells);
/* InfMatterSPBasis(1, density, tzMax = 1, nShells, nParticleShells): pure
 * neutron matter, Minnesota potential. */
orc_t *orc_create_infmatter(double density, int nShells, int nParticleShells);
void orc_destroy(orc_t *o);

/* Build every V chain (reference loadVnn_*, src/main.cpp:293-366) and the
 * eight T2 maps (loadMaps, src/load.cpp:389-532). */
int orc_load(orc_t *o);
/* Before orc_load: keep V^pppp out of memory and rebuild each channel's block
 * inside orc_ladder (the reference's -m 1 ladder). */
void orc_set_matrix_free(orc_t *o, int on);

void orc_set_threads(int n);

size_t orc_nsp(const orc_t *o);
size_t orc_nholes(const orc_t *o);
size_t orc_nchannels(const orc_t *o);
size_t orc_nt2(const orc_t *o);
double orc_eref(const orc_t *o);
void orc_get_eps(const orc_t *o, double *eps);
/* Quantum numbers, 5 ints per state (HEG nx ny nz sz tz; pairing level spin 0 0 0). */
void orc_get_qn(const orc_t *o, int32_t *qn);
/* Per-channel (permuted order) direct pp/hh and mod ph/hp dims. */
void orc_get_dims(const orc_t *o, int64_t *pp, int64_t *hh, int64_t *mph, int64_t *mhp);
/* Labels (a,b,i,j) of every T2 element in pphh layout. */
void orc_t2_labels(const orc_t *o, int32_t *a, int32_t *b, int32_t *i, int32_t *j);
/* Labels ((a,i),(j,b)) of every element of the phhp (mod) layout. */
void orc_tph_labels(const orc_t *o, int32_t *a, int32_t *i, int32_t *j, int32_t *b);
/* Flat V^hhpp in hhpp layout (hh x pp blocks, src/load.cpp:154-174). */
void orc_get_vhhpp(const orc_t *o, double *v);

/* MBPT2 correlation energy (src/MBPT2Corr.cpp:6-38). */
double orc_mbpt2(const orc_t *o);

/* Diagram pieces on caller-supplied amplitudes (all in oracle layouts). */
/* calcT2Diff_pphh: D (pphh) from T (pphh)       src/calc.cpp:132-231 */
void orc_ladder(const orc_t *o, const double *t, double *d);
/* calcT2Diff_phhp: R (phhp) from T^ph (phhp)    src/calc.cpp:264-339 */
void orc_ring(const orc_t *o, const double *tph, double *r);
/* mapToDistributed(ai_jb): T (pphh) -> T^ph     src/Chain.hpp:757-802 */
void orc_scatter_ph(const orc_t *o, const double *t, double *tph);
/* The scalar one-body intermediates of calcT2Diff_h2pp_h1 / _p2_p1hh
 * (the `value` of src/calc.cpp:81-84 and :15-20): chi[nh], psi[nsp-nh]. */
void orc_onebody(const orc_t *o, const double *t, double *chi, double *psi);
/* One full iteration from T_old (src/main.cpp:589-670): writes T_new and
 * returns the correlation energy of T_new. */
double orc_step(orc_t *o, const double *t_old, double *t_new, double mixing);

/* The whole CCD solve with the reference protocol (src/main.cpp:491-682):
 * T2 = 0, mixing 0.5, E_prev = 100, loop while |dE| > tol, halt after
 * iteration 100.  energies[k] = corrCCD after iteration k (k < cap).
 * Returns the number of iterations; *stop = 0 converged, 1 max-iterations,
 * 2 left the loop on a NaN energy difference. */
int orc_solve(orc_t *o, double tol, double *energies, int cap, int *stop,
              double *seconds);
/* Final amplitudes of the last solve (pphh layout). */
void orc_get_t2(const orc_t *o, double *t);

#ifdef __cplusplus
}
#endif
#endif
