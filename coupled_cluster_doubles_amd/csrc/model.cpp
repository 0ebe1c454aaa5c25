// Host problem model: the single-particle bases, channel blocking and
// interaction of the reference code, restated so the library can run the
// reference's inputs on a machine that only has this repository.
//
// Behaviour follows (file:line in justinlietz/coupled_cluster_doubles):
//   src/PairingSPBasis.cpp:14-58, 127-181        pairing basis, <pq||rs>, channels
//   src/ElectronGasSPBasis.cpp:14-210, 424-624   electron gas basis, Coulomb, channels
//   src/InfMatterSPBasis.cpp:14-184, 355-466     pure neutron matter, Minnesota potential
//   src/SPBasis.cpp:7-40                         E_ref and normal ordering
//   src/SPBasis.cpp:80-819                       channel dims and sorted pair maps
//   src/main.cpp:94-115 + src/balance.cpp:91-144 channel order at one rank
//   src/load.cpp:112-318                         V blocks (host copy, optional)
// Arithmetic that produces printed numbers (constants, energies, matrix
// elements) is evaluated in the reference's operation order so the values
// are bitwise identical; this translation unit is compiled with
// -ffp-contract=off for that reason.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "ccd_amd.h"
#include "internal.hpp"

namespace {

// n is a sum of three squares (Legendre), as the shell counter uses it
// (src/util.cpp:1-9).
bool three_squares(unsigned n) {
    if (n == 0) return true;
    while (n % 4 == 0) n /= 4;
    return (n % 8) != 7;
}

// Smallest r with r^2 >= x, as src/util.cpp:11-19 computes it.
unsigned sqrt_at_least(unsigned x) {
    unsigned r = 1;
    while (r * r < x) ++r;
    return r;
}

// The (k-1)-th positive integer that is a sum of three squares, i.e. the
// largest n^2 kept when k shells (n^2 = 0 included) are requested
// (src/ElectronGasSPBasis.cpp:38-43 and :69-74).
unsigned shell_energy(int k) {
    unsigned e = 0;
    for (int found = 1; found < k;) {
        ++e;
        if (three_squares(e)) ++found;
    }
    return e;
}

}  // namespace

constexpr int kInfMatter = CCD_GEN_INF_MATTER;

struct ccd_model {
    int kind = 0;                    // CCD_GEN_PAIRING, CCD_GEN_ELECTRON_GAS or CCD_GEN_INF_MATTER
    int32_t nsp = 0, nh = 0;
    std::vector<int32_t> qn;         // 5 per state
    std::vector<double> eps;         // normal ordered
    double eref = 0.0;
    double g = 0.0, xi = 0.0, r_s = 0.0, L = 0.0, e_sq = 0.0;
    double density = 0.0, inf_L = 0.0;
    int32_t nmax = 0;
    std::vector<double> table;       // electron gas direct term per momentum transfer;
                                     // infinite matter: the 3 Minnesota prefactors
    int32_t nch = 0;
    std::vector<int32_t> pp_dim, hh_dim, ph_dim, hp_dim;
    std::vector<int32_t> pp_pairs, hh_pairs, ph_pairs, hp_pairs;
    std::vector<double> v_hhhh, v_hhpp, v_pppp, v_modA, v_modB;
    bool host_v = false;
    ccd_problem prob{};
    ccd_problem prob_host{};

    const int32_t *q(int p) const { return qn.data() + 5 * (size_t)p; }

    // Direct (non-antisymmetrised) Coulomb element <pq|v|rs>
    // (src/ElectronGasSPBasis.cpp:424-471); momentum transfer from the table.
    double heg_direct(int p, int qq, int r, int s) const {
        const int32_t *a = q(p), *b = q(qq), *c = q(r), *d = q(s);
        if (a[3] != c[3] || b[3] != d[3]) return 0.0;
        if (a[0] + b[0] != c[0] + d[0] || a[1] + b[1] != c[1] + d[1] || a[2] + b[2] != c[2] + d[2])
            return 0.0;
        const int m = nmax, w = 4 * nmax + 1;
        size_t idx = ((size_t)(a[0] - c[0] + 2 * m) * w + (size_t)(a[1] - c[1] + 2 * m)) * w +
                     (size_t)(a[2] - c[2] + 2 * m);
        return table[idx];
    }

    // Minnesota potential <pq|v|rs>, not antisymmetrised
    // (src/InfMatterSPBasis.cpp:361-449), in its operation order.
    double minnesota(int p, int qq, int r, int s) const {
        const int32_t *qi = q(p), *qj = q(qq), *qk = q(r), *ql = q(s);
        if (qi[0] + qj[0] != qk[0] + ql[0] || qi[1] + qj[1] != qk[1] + ql[1] || qi[2] + qj[2] != qk[2] + ql[2])
            return 0.0;
        const double Lb = inf_L;
        const double V_0R = 200, V_0T = 178, V_0S = 91.85;
        const double kappa_R = 1.487, kappa_T = 0.639, kappa_S = 0.465;
        double qSquared = 0.;
        for (int i = 0; i < 3; i++) {
            const double ki = 2 * M_PI * qi[i] / Lb, kj = 2 * M_PI * qj[i] / Lb;
            const double kk = 2 * M_PI * qk[i] / Lb, kl = 2 * M_PI * ql[i] / Lb;
            const double bra = 0.5 * (ki - kj), ket = 0.5 * (kk - kl), tr = bra - ket;
            qSquared += tr * tr;
        }
        const double V_R = V_0R / (Lb * Lb * Lb) * std::pow(M_PI / kappa_R, 1.5) * std::exp(-qSquared / (4 * kappa_R));
        const double V_T = -V_0T / (Lb * Lb * Lb) * std::pow(M_PI / kappa_T, 1.5) * std::exp(-qSquared / (4 * kappa_T));
        const double V_S = -V_0S / (Lb * Lb * Lb) * std::pow(M_PI / kappa_S, 1.5) * std::exp(-qSquared / (4 * kappa_S));
        auto kron = [](int a, int b) { return a != b ? 0 : 1; };
        auto exch = [](int i, int j, int k, int l) { return (i == l && j == k) ? 1 : 0; };
        const double spinEx = exch(qi[3], qj[3], qk[3], ql[3]);
        const double isoSpinEx = exch(qi[4], qj[4], qk[4], ql[4]);
        const double IsIt = kron(qi[3], qk[3]) * kron(qj[3], ql[3]) * kron(qi[4], qk[4]) * kron(qj[4], ql[4]);
        const double PsIt = spinEx * kron(qi[4], qk[4]) * kron(qj[4], ql[4]);
        const double PsPt = spinEx * isoSpinEx;
        const double IsPt = kron(qi[3], qk[3]) * kron(qj[3], ql[3]) * isoSpinEx;
        return 0.5 * (V_R + 0.5 * V_T + 0.5 * V_S) * IsIt + 0.25 * (V_T - V_S) * PsIt -
               0.5 * (V_R + 0.5 * V_T + 0.5 * V_S) * PsPt - 0.25 * (V_T - V_S) * IsPt;
    }

    // Antisymmetrised <pq||rs> (calcTBME of each basis).
    double tbme(int p, int qq, int r, int s) const {
        if (kind == kInfMatter) return minnesota(p, qq, r, s) - minnesota(p, qq, s, r);  // :355-359
        if (kind == CCD_GEN_PAIRING) {  // src/PairingSPBasis.cpp:127-141
            const int32_t *a = q(p), *b = q(qq), *c = q(r), *d = q(s);
            double v = 0.0;
            if (a[0] == b[0] && c[0] == d[0] && a[1] == -b[1] && -c[1] == d[1]) {
                v = -0.5 * g;
                if (a[1] == -c[1]) v = -v;
            }
            return v;
        }
        return heg_direct(p, qq, r, s) - heg_direct(p, qq, s, r);  // :419-422
    }

    // Raw channel keys (index functions before the balancer's permutation).
    size_t raw_direct(int p, int qq) const {
        const int32_t *a = q(p), *b = q(qq);
        if (kind == CCD_GEN_PAIRING) return (size_t)((a[1] + b[1] + 2) / 2);
        const size_t w = 4 * (size_t)nmax + 1, h = w / 2;
        return ((size_t)(a[0] + b[0] + (long)h) * w * w * 3) + ((size_t)(a[1] + b[1] + (long)h) * w * 3) +
               ((size_t)(a[2] + b[2] + (long)h) * 3) + (size_t)((a[3] + b[3] + 2) / 2);
    }
    size_t raw_mod(int p, int qq) const {
        const int32_t *a = q(p), *b = q(qq);
        if (kind == CCD_GEN_PAIRING) return (size_t)((a[1] - b[1] + 2) / 2);
        const size_t w = 4 * (size_t)nmax + 1, h = w / 2;
        return ((size_t)(a[0] - b[0] + (long)h) * w * w * 3) + ((size_t)(a[1] - b[1] + (long)h) * w * 3) +
               ((size_t)(a[2] - b[2] + (long)h) * 3) + (size_t)((a[3] - b[3] + 2) / 2);
    }

    // Plane-wave momenta with n^2 up to the n_shells-th shell, two spin states,
    // one isospin state (tz = -1), ordered by n^2, then nx, ny, nz ascending,
    // then sz = -1, +1 (src/ElectronGasSPBasis.cpp:38-170; the same loops in
    // src/InfMatterSPBasis.cpp:14-163 for pure neutron matter).  Holes: the
    // states of the first n_particle_shells shells.
    void plane_wave_states(int n_shells, int n_particle_shells) {
        const unsigned emax = shell_energy(n_shells), ef = shell_energy(n_particle_shells);
        const int range = (int)sqrt_at_least(emax);
        std::vector<std::vector<int32_t>> by_shell(emax + 1);
        int nm = 0;
        for (int x = -range; x <= range; x++)
            for (int y = -range; y <= range; y++)
                for (int z = -range; z <= range; z++) {
                    unsigned e = (unsigned)(x * x + y * y + z * z);
                    if (e > emax) continue;
                    nm = std::max(nm, x);
                    for (int s = -1; s <= 1; s += 2) {
                        auto &v = by_shell[e];
                        v.insert(v.end(), {x, y, z, s, -1});
                    }
                }
        for (auto &v : by_shell) qn.insert(qn.end(), v.begin(), v.end());
        nsp = (int32_t)(qn.size() / 5);
        nmax = nm;
        nh = 0;
        for (int p = 0; p < nsp; p++) {
            const int32_t *a = q(p);
            if ((unsigned)(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) <= ef) nh++;
        }
    }

    void reference_energy_and_normal_order() {  // src/SPBasis.cpp:7-40
        double e = 0.0;
        for (int i = 0; i < nh; i++)
            for (int j = 0; j < nh; j++) e += 0.5 * tbme(i, j, i, j);
        for (int i = 0; i < nh; i++) e += eps[i];
        eref = e;
        for (int p = 0; p < nsp; p++) {
            double ep = eps[p];
            for (int j = 0; j < nh; j++) ep += tbme(p, j, p, j);
            eps[p] = ep;
        }
    }

    void build_channels(size_t nraw) {
        std::vector<int64_t> pp(nraw), hh(nraw), hp(nraw), ph(nraw), mph(nraw), mhp(nraw);
        for (int p = 0; p < nsp; p++)
            for (int r = 0; r < nsp; r++) {
                size_t c = raw_direct(p, r), m = raw_mod(p, r);
                bool hp_ = p < nh, hq = r < nh;
                if (hp_ && hq) hh[c]++;
                else if (hp_) { hp[c]++; mhp[m]++; }
                else if (hq) { ph[c]++; mph[m]++; }
                else pp[c]++;
            }
        // Cost model of src/main.cpp:94-112 (direct dims) and first-fit
        // decreasing at one rank: descending cost, ties by raw index.
        std::vector<double> cost(nraw);
        for (size_t c = 0; c < nraw; c++) {
            double P = (double)pp[c], H = (double)hh[c], PH = (double)ph[c], HP = (double)hp[c];
            double x = 0;
            x += H * H; x += H * HP; x += H * P; x += HP * PH;
            x += HP * P; x += P * P; x += HP * PH; x += HP * HP;
            cost[c] = x;
        }
        std::vector<size_t> order(nraw);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return cost[a] > cost[b]; });
        std::vector<int32_t> inv(nraw);
        for (size_t k = 0; k < nraw; k++) inv[order[k]] = (int32_t)k;
        nch = (int32_t)nraw;
        pp_dim.resize(nraw); hh_dim.resize(nraw); ph_dim.resize(nraw); hp_dim.resize(nraw);
        for (size_t k = 0; k < nraw; k++) {
            size_t r = order[k];
            pp_dim[k] = (int32_t)pp[r]; hh_dim[k] = (int32_t)hh[r];
            ph_dim[k] = (int32_t)mph[r]; hp_dim[k] = (int32_t)mhp[r];
        }
        auto prefix = [&](const std::vector<int32_t> &d) {
            std::vector<int64_t> o(d.size() + 1, 0);
            for (size_t k = 0; k < d.size(); k++) o[k + 1] = o[k] + d[k];
            return o;
        };
        auto opp = prefix(pp_dim), ohh = prefix(hh_dim), oph = prefix(ph_dim), ohp = prefix(hp_dim);
        pp_pairs.resize(2 * opp.back()); hh_pairs.resize(2 * ohh.back());
        ph_pairs.resize(2 * oph.back()); hp_pairs.resize(2 * ohp.back());
        // Lexicographic enumeration = the std::sort order of src/SPBasis.cpp:372-378.
        for (int p = 0; p < nsp; p++)
            for (int r = 0; r < nsp; r++) {
                int32_t c = inv[raw_direct(p, r)], m = inv[raw_mod(p, r)];
                int32_t *dst;
                if (p < nh && r < nh) dst = &hh_pairs[2 * ohh[c]++];
                else if (p >= nh && r >= nh) dst = &pp_pairs[2 * opp[c]++];
                else if (p >= nh) dst = &ph_pairs[2 * oph[m]++];
                else dst = &hp_pairs[2 * ohp[m]++];
                dst[0] = p; dst[1] = r;
            }
    }

    void fill_problem() {
        prob = ccd_problem{};
        prob.n_sp = nsp; prob.n_holes = nh; prob.eps = eps.data();
        prob.n_chan = nch; prob.pp_dim = pp_dim.data(); prob.hh_dim = hh_dim.data();
        prob.pp_pairs = pp_pairs.data(); prob.hh_pairs = hh_pairs.data();
        prob.n_mod = nch; prob.ph_dim = ph_dim.data(); prob.hp_dim = hp_dim.data();
        prob.ph_pairs = ph_pairs.data(); prob.hp_pairs = hp_pairs.data();
        prob.gen_kind = kind; prob.qn = qn.data(); prob.gen_g = kind == kInfMatter ? inf_L : g;
        prob.gen_nmax = nmax; prob.gen_table = table.empty() ? nullptr : table.data();
    }

    // Host V chains in the reference layouts (src/load.cpp:112-318).
    void build_host_v() {
        if (host_v) return;
        size_t n_hhhh = 0, n_hhpp = 0, n_pppp = 0, n_a = 0, n_b = 0;
        for (int c = 0; c < nch; c++) {
            n_hhhh += (size_t)hh_dim[c] * hh_dim[c];
            n_hhpp += (size_t)hh_dim[c] * pp_dim[c];
            n_pppp += (size_t)pp_dim[c] * pp_dim[c];
            n_a += (size_t)hp_dim[c] * ph_dim[c];
            n_b += (size_t)hp_dim[c] * hp_dim[c];
        }
        v_hhhh.assign(n_hhhh, 0.0); v_hhpp.assign(n_hhpp, 0.0); v_pppp.assign(n_pppp, 0.0);
        v_modA.assign(n_a, 0.0); v_modB.assign(n_b, 0.0);
        size_t o1 = 0, o2 = 0, o3 = 0, o4 = 0, o5 = 0, ip = 0, ih = 0, iph = 0, ihp = 0;
        for (int c = 0; c < nch; c++) {
            const int P = pp_dim[c], H = hh_dim[c];
            const int32_t *ppl = &pp_pairs[2 * ip], *hhl = &hh_pairs[2 * ih];
            for (int k = 0; k < H; k++)
                for (int l = 0; l < H; l++)
                    v_hhhh[o1++] = tbme(hhl[2 * k], hhl[2 * k + 1], hhl[2 * l], hhl[2 * l + 1]);
            for (int k = 0; k < H; k++)
                for (int l = 0; l < P; l++)
                    v_hhpp[o2++] = tbme(hhl[2 * k], hhl[2 * k + 1], ppl[2 * l], ppl[2 * l + 1]);
            for (int k = 0; k < P; k++)
                for (int l = 0; l < P; l++)
                    v_pppp[o3++] = tbme(ppl[2 * k], ppl[2 * k + 1], ppl[2 * l], ppl[2 * l + 1]);
            ip += P; ih += H;
            const int MP = ph_dim[c], MH = hp_dim[c];
            const int32_t *phl = &ph_pairs[2 * iph], *hpl = &hp_pairs[2 * ihp];
            for (int k = 0; k < MH; k++)
                for (int l = 0; l < MP; l++) {  // (i,a) x (b,j): <ij||ab>
                    int i = hpl[2 * k], a = hpl[2 * k + 1], b = phl[2 * l], j = phl[2 * l + 1];
                    v_modA[o4++] = tbme(i, j, a, b);
                }
            for (int k = 0; k < MH; k++)
                for (int l = 0; l < MH; l++) {  // (i,b) x (j,a): <ia||bj>
                    int i = hpl[2 * k], b = hpl[2 * k + 1], j = hpl[2 * l], a = hpl[2 * l + 1];
                    v_modB[o5++] = tbme(i, a, b, j);
                }
            iph += MP; ihp += MH;
        }
        host_v = true;
        prob_host = prob;
        prob_host.gen_kind = CCD_GEN_NONE;
        prob_host.v_hhhh = v_hhhh.data(); prob_host.v_hhpp = v_hhpp.data();
        prob_host.v_pppp = v_pppp.data(); prob_host.v_modA = v_modA.data();
        prob_host.v_modB = v_modB.data();
    }
};

extern "C" {

int ccd_model_pairing(int n_pair_states, int n_particles, double g, double xi, ccd_model **out) {
    ccd::clear_error();
    if (!out || n_pair_states <= 0 || n_particles < 0 || n_particles > 2 * n_pair_states)
        return ccd::fail(CCD_E_INVALID, "ccd_model_pairing: bad arguments");
    ccd_model *m = new (std::nothrow) ccd_model();
    if (!m) return ccd::fail(CCD_E_NOMEM, "ccd_model_pairing: out of memory");
    m->kind = CCD_GEN_PAIRING;
    m->g = g; m->xi = xi;
    m->nsp = 2 * n_pair_states; m->nh = n_particles;
    m->qn.assign(5 * (size_t)m->nsp, 0);
    m->eps.resize(m->nsp);
    for (int l = 0; l < n_pair_states; l++) {  // src/PairingSPBasis.cpp:27-44
        m->qn[5 * (2 * l)] = l + 1; m->qn[5 * (2 * l) + 1] = +1;
        m->qn[5 * (2 * l + 1)] = l + 1; m->qn[5 * (2 * l + 1) + 1] = -1;
    }
    for (int p = 0; p < m->nsp; p++) m->eps[p] = (m->qn[5 * p] - 1.0) * xi;  // :48-58
    m->reference_energy_and_normal_order();
    m->build_channels(3);
    m->fill_problem();
    *out = m;
    return CCD_OK;
}

int ccd_model_electron_gas(double r_s, int n_shells, int n_particle_shells, ccd_model **out) {
    ccd::clear_error();
    if (!out || n_shells < 1 || n_particle_shells < 1 || n_particle_shells > n_shells || !(r_s > 0))
        return ccd::fail(CCD_E_INVALID, "ccd_model_electron_gas: bad arguments");
    ccd_model *m = new (std::nothrow) ccd_model();
    if (!m) return ccd::fail(CCD_E_NOMEM, "ccd_model_electron_gas: out of memory");
    m->kind = CCD_GEN_ELECTRON_GAS;
    m->r_s = r_s;
    m->plane_wave_states(n_shells, n_particle_shells);
    const int nmax = m->nmax, nh = m->nh;
    // Constants and kinetic energies of src/ElectronGasSPBasis.cpp:178-207.
    double hbarc = 0.1973269788;
    double ev_per_hartree = 27.21138505;
    hbarc = hbarc * 10000 / ev_per_hartree;
    double massc2 = 0.5109989461;
    massc2 = massc2 * 1000000 / ev_per_hartree;
    double alpha = 1.0 / 137.035999139;
    double r_bohr = hbarc / (massc2 * alpha);
    double r_0 = r_s * r_bohr;
    double kin = hbarc * hbarc / (2. * massc2);
    m->L = std::pow((4. / 3.) * M_PI * nh, 1. / 3.) * r_0;
    m->e_sq = alpha * hbarc;
    m->eps.resize(m->nsp);
    for (int p = 0; p < m->nsp; p++) {
        const int32_t *a = m->q(p);
        m->eps[p] = (kin * 4 * M_PI * M_PI / (m->L * m->L)) * (a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    }
    // Direct Coulomb term per momentum transfer, evaluated exactly as
    // src/ElectronGasSPBasis.cpp:447-461 does per element.
    const int w = 4 * nmax + 1;
    m->table.assign((size_t)w * w * w, 0.0);
    const double pref = m->e_sq / (m->L * m->L * m->L);
    const double mu = 0.0;
    for (int dx = -2 * nmax; dx <= 2 * nmax; dx++)
        for (int dy = -2 * nmax; dy <= 2 * nmax; dy++)
            for (int dz = -2 * nmax; dz <= 2 * nmax; dz++) {
                const int dd[3] = {dx, dy, dz};
                double rel = 0.0;
                for (int k = 0; k < 3; k++) {
                    double term = 2 * M_PI * dd[k] / m->L;
                    rel += term * term;
                }
                double v = (rel < 1.e-5) ? 0.0 : pref * 4 * M_PI / (rel + mu * mu);
                m->table[((size_t)(dx + 2 * nmax) * w + (dy + 2 * nmax)) * w + (dz + 2 * nmax)] = v;
            }
    m->reference_energy_and_normal_order();
    m->build_channels((size_t)w * w * w * 3);
    m->fill_problem();
    *out = m;
    return CCD_OK;
}

int ccd_model_inf_matter(double density, int tz_max, int n_shells, int n_particle_shells, ccd_model **out) {
    ccd::clear_error();
    if (!out || n_shells < 1 || n_particle_shells < 1 || n_particle_shells > n_shells || !(density > 0))
        return ccd::fail(CCD_E_INVALID, "ccd_model_inf_matter: bad arguments");
    if (tz_max != 1)
        return ccd::fail(CCD_E_INVALID, "ccd_model_inf_matter: only pure neutron matter (tz_max = 1) is supported");
    ccd_model *m = new (std::nothrow) ccd_model();
    if (!m) return ccd::fail(CCD_E_NOMEM, "ccd_model_inf_matter: out of memory");
    m->kind = kInfMatter;
    m->density = density;
    m->plane_wave_states(n_shells, n_particle_shells);
    // Kinetic energies of src/InfMatterSPBasis.cpp:166-184 (neutron mass);
    // calc_TBME_not_antisym recomputes the same box length (:381).
    const double hbarc = 197.3269788, massc2 = 939.565;
    const double prefactor = hbarc * hbarc / (2. * massc2);
    m->L = std::pow((size_t)m->nh / density, 1. / 3.);
    m->inf_L = std::pow((size_t)m->nh / m->density, 1. / 3.);
    m->eps.resize(m->nsp);
    for (int p = 0; p < m->nsp; p++) {
        const int32_t *a = m->q(p);
        m->eps[p] = (prefactor * 4 * M_PI * M_PI / (m->L * m->L)) * (a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    }
    // channels keyed like the electron gas: (4 nMax + 1)^3 x 3 x 1
    // (src/InfMatterSPBasis.cpp:196-240 with chanTzMax = -2)
    const int w = 4 * m->nmax + 1;
    m->reference_energy_and_normal_order();
    m->build_channels((size_t)w * w * w * 3);
    // the Minnesota prefactors of src/InfMatterSPBasis.cpp:412-414, for the
    // on-device generator (the Gaussians are evaluated per element there)
    {
        const double Lb = m->inf_L;
        const double V_0R = 200, V_0T = 178, V_0S = 91.85;
        const double kappa_R = 1.487, kappa_T = 0.639, kappa_S = 0.465;
        m->table = {V_0R / (Lb * Lb * Lb) * std::pow(M_PI / kappa_R, 1.5),
                    -V_0T / (Lb * Lb * Lb) * std::pow(M_PI / kappa_T, 1.5),
                    -V_0S / (Lb * Lb * Lb) * std::pow(M_PI / kappa_S, 1.5)};
    }
    m->fill_problem();
    *out = m;
    return CCD_OK;
}

int ccd_model_destroy(ccd_model *m) {
    delete m;
    return CCD_OK;
}

int ccd_model_eref(const ccd_model *m, double *eref) {
    if (!m || !eref) return ccd::fail(CCD_E_INVALID, "ccd_model_eref: null argument");
    *eref = m->eref;
    return CCD_OK;
}

int ccd_model_sizes(const ccd_model *m, int32_t *n_sp, int32_t *n_holes, int32_t *n_chan, int64_t *n_t2) {
    if (!m) return ccd::fail(CCD_E_INVALID, "ccd_model_sizes: null model");
    if (n_sp) *n_sp = m->nsp;
    if (n_holes) *n_holes = m->nh;
    if (n_chan) *n_chan = m->nch;
    if (n_t2) {
        int64_t s = 0;
        for (int c = 0; c < m->nch; c++) s += (int64_t)m->pp_dim[c] * m->hh_dim[c];
        *n_t2 = s;
    }
    return CCD_OK;
}

int ccd_model_problem(ccd_model *m, const ccd_problem **out) {
    if (!m || !out) return ccd::fail(CCD_E_INVALID, "ccd_model_problem: null argument");
    *out = &m->prob;
    return CCD_OK;
}

int ccd_model_problem_host_v(ccd_model *m, const ccd_problem **out) {
    if (!m || !out) return ccd::fail(CCD_E_INVALID, "ccd_model_problem_host_v: null argument");
    try {
        m->build_host_v();
    } catch (const std::bad_alloc &) {
        return ccd::fail(CCD_E_NOMEM, "ccd_model_problem_host_v: out of host memory");
    }
    *out = &m->prob_host;
    return CCD_OK;
}

}  // extern "C"
