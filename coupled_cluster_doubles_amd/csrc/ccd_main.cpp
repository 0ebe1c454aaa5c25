// ccd -- drop-in driver for the reference's ccd.exe on one MI355X.
//
// Accepts the reference's options (src/parse.cpp:131-155, defaults
// :170-179, -c FILE re-parsing :209-229) and prints the reference's report
// lines (src/main.cpp:32, 482-487, 512-519, 672-700) with %.17g energies.
// The basis is built by the host model of libccd_amd; MBPT2 and every CCD
// iteration run on the GPU through the C ABI, driven by the reference's
// loop protocol (T2 = 0, mixing 0.5, E_prev = 100, halt after iteration 100).
#include <argp.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ccd_amd.h"

namespace {

struct Options {
    int basis = 1;  // 0 pairing, 1 infMatter, 2 electronGas, 3 qDots (parse.cpp:26-29)
    std::string config;
    double density = 0.08, g = 0.5, r_s = 0.5, xi = 1.0, tolerance = 1.e-8;
    int nParticleShells = 2, nParticles = 4, nPairStates = 4, nShells = 4, tzMax = 1;
    int saveMemory = 0, verbose = 0;
    int precision = CCD_PREC_FP64;
};

enum {
    k_pairing = 0x100, k_infMatter = 0x101, k_electronGas = 0x102, k_qDots = 0x103,
    k_precision = 0x200,
};

argp_option options[] = {
    {"config", 'c', "FILE", 0, "read option defaults from FILE (one option per line)", 0},
    {"saveMemory", 'm', "INT", 0, "1: rebuild V^pppp on the device every iteration instead of storing it; 0: store it", 0},
    {"time", 'T', "FILE", 0, "accepted for compatibility; one GPU needs no load-balancer timings", 0},
    {"timeOverwrite", 'W', 0, 0, "accepted for compatibility; no effect", 0},
    {"model", 'M', "FILE", 0, "accepted for compatibility; no load-balancer model is read", 0},
    {"verbose", 'v', "INT", 0, "how much to print (0 = the report lines only)", 0},
    {"precision", k_precision, "fp64|fp32", 0, "arithmetic of the contractions (default fp64)", 0},
    {0, 0, 0, 0, "Which basis:", 0},
    {"pairing", k_pairing, 0, 0, "doubly degenerate equally spaced levels with constant pairing", 0},
    {"infMatter", k_infMatter, 0, 0, "nuclear matter in a periodic box, Minnesota interaction", 0},
    {"qDots", k_qDots, 0, 0, "quantum dots (not available)", 0},
    {"electronGas", k_electronGas, 0, 0, "homogeneous electron gas in a periodic box, Coulomb interaction", 0},
    {0, 0, 0, 0, "Basis parameters:", 0},
    {"density", 'd', "FLOAT", 0, "nuclear matter: density in fm^-3", 0},
    {"g", 'g', "FLOAT", 0, "pairing model: strength of the pairing term", 0},
    {"nParticleShells", 'p', "INT", 0, "number of filled momentum shells (sets the particle count)", 0},
    {"nParticles", 'n', "INT", 0, "pairing model: particle count", 0},
    {"nPairStates", 'l', "INT", 0, "pairing model: number of levels", 0},
    {"r_s", 'r', "FLOAT", 0, "electron gas: Wigner-Seitz radius in Bohr", 0},
    {"nShells", 's', "INT", 0, "number of momentum shells in the basis", 0},
    {"tolerance", 't', "FLOAT", 0, "stop once the correlation energy changes by at most this much", 0},
    {"tzMax", 'z', "INT", 0, "nuclear matter: 1 = neutrons only, 2 = neutrons and protons", 0},
    {"xi", 'x', "FLOAT", 0, "pairing model: spacing of the levels", 0},
    {0, 0, 0, 0, 0, 0},
};

error_t parse_opt(int key, char *arg, argp_state *state) {
    Options *o = static_cast<Options *>(state->input);
    switch (key) {
        case k_pairing: case k_infMatter: case k_electronGas: case k_qDots:
            o->basis = key - 0x100; break;
        case 'c': o->config = arg; break;
        case 'd': o->density = atof(arg); break;
        case 'g': o->g = atof(arg); break;
        case 'p': o->nParticleShells = atoi(arg); break;
        case 'n': o->nParticles = atoi(arg); break;
        case 'l': o->nPairStates = atoi(arg); break;
        case 'r': o->r_s = atof(arg); break;
        case 's': o->nShells = atoi(arg); break;
        case 't': o->tolerance = atof(arg); break;
        case 'z': o->tzMax = atoi(arg); break;
        case 'x': o->xi = atof(arg); break;
        case 'm': o->saveMemory = atoi(arg); break;
        case 'T': case 'W': case 'M': break;  // load-balancer options: nothing to balance on one GPU
        case 'v': o->verbose = atoi(arg); break;
        case k_precision:
            if (!strcmp(arg, "fp64")) o->precision = CCD_PREC_FP64;
            else if (!strcmp(arg, "fp32")) o->precision = CCD_PREC_FP32;
            else argp_error(state, "unknown precision '%s'", arg);
            break;
        default: return ARGP_ERR_UNKNOWN;
    }
    return 0;
}

argp parser = {options, parse_opt, "", "CCD on one MI355X (drop-in for ccd.exe)", 0, 0, 0};

int die(const char *what, int rc) {
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, ccd_last_error());
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    Options o;
    argp_parse(&parser, argc, argv, 0, 0, &o);
    if (!o.config.empty()) {  // config tokens first, then the command line again
        std::ifstream f(o.config);
        if (!f) { fprintf(stderr, "cannot open config file %s\n", o.config.c_str()); return 1; }
        std::stringstream ss;
        ss << f.rdbuf();
        std::vector<std::string> toks;
        std::string tok;
        while (ss >> tok) toks.push_back(tok);
        std::vector<char *> cargv;
        cargv.push_back(argv[0]);
        for (auto &t : toks) cargv.push_back(&t[0]);
        cargv.push_back(nullptr);
        std::string keep = o.config;
        argp_parse(&parser, (int)cargv.size() - 1, cargv.data(), 0, 0, &o);
        argp_parse(&parser, argc, argv, 0, 0, &o);
        o.config = keep;
    }
    printf("rank: %d size: %d\n", 0, 1);
    auto t_total = std::chrono::steady_clock::now();

    ccd_model *model = nullptr;
    int rc;
    switch (o.basis) {
        case 0: rc = ccd_model_pairing(o.nPairStates, o.nParticles, o.g, o.xi, &model); break;
        case 2: rc = ccd_model_electron_gas(o.r_s, o.nShells, o.nParticleShells, &model); break;
        case 1: rc = ccd_model_inf_matter(o.density, o.tzMax, o.nShells, o.nParticleShells, &model); break;
        default:
            printf("not ready\n");
            return 1;
    }
    if (rc) return die("basis setup", rc);
    int32_t nsp = 0, nh = 0, nch = 0;
    int64_t nt2 = 0;
    ccd_model_sizes(model, &nsp, &nh, &nch, &nt2);
    if (o.verbose >= 1) {
        if (o.basis == 0) {
            printf("Basis 0: Pairing\n");
            printf("xi = %f, g = %f, nPairStates = %d, nParticles = %d, nSpstates = %d\n", o.xi, o.g,
                   o.nPairStates, o.nParticles, nsp);
        } else if (o.basis == 1) {
            printf("Basis 1: infMatter\n");
            printf("density = %f, tzMax = %d, nShells = %d, nParticleShells = %d, nSpStates = %d, nParticles = %d\n",
                   o.density, o.tzMax, o.nShells, o.nParticleShells, nsp, nh);
        } else {
            printf("Basis 2: electronGas\n");
            printf("r_s = %f, nShells = %d, nParticleShells = %d, nSpStates = %d, nParticles = %d\n", o.r_s,
                   o.nShells, o.nParticleShells, nsp, nh);
        }
        printf("nChannels: %d\n", nch);
    }
    double eref = 0.0;
    ccd_model_eref(model, &eref);
    const ccd_problem *prob = nullptr;
    ccd_model_problem(model, &prob);
    ccd_ctx *ctx = nullptr;
    // -m 1 regenerates V^pppp inside the ladder
    const int mode = o.precision | (o.saveMemory ? CCD_MATRIX_FREE : 0);
    if ((rc = ccd_ctx_create(prob, mode, &ctx))) return die("ccd_ctx_create", rc);
    // the fastest launch schedule for this problem (same results under any)
    if ((rc = ccd_set_schedule(ctx, CCD_SCHED_AUTO))) return die("ccd_set_schedule", rc);
    int64_t dbytes = 0;
    ccd_ctx_info(ctx, nullptr, &dbytes);
    printf("rank %d device arrays require %g GB of memory\n", 0, dbytes / 1.e9);

    auto t0 = std::chrono::steady_clock::now();
    double mbpt2 = 0.0;
    if ((rc = ccd_mbpt2(ctx, &mbpt2))) return die("ccd_mbpt2", rc);
    auto t1 = std::chrono::steady_clock::now();
    printf("mbpt loops time: %f\n", std::chrono::duration<double>(t1 - t0).count());
    const double A = (double)nh;
    printf("Reference Energy: %.17g\n", eref);
    printf("Reference Energy/A: %.17g\n", eref / A);
    printf("MBPT2Corr: %.17g\n", mbpt2);
    printf("MBPT2Corr/A: %.17g\n", mbpt2 / A);
    printf("E_mbpt2: %.17g\n", eref + mbpt2);
    printf("E_mbpt2/A: %.17g\n", (eref + mbpt2) / A);

    // src/main.cpp:491-682
    double energyDiff = 1.0, corrCCD = 100.0, prev = 100.0;
    unsigned long iterations = 0;
    const double mixing = 0.5;
    const unsigned long max_iters = 100;
    if ((rc = ccd_reset(ctx))) return die("ccd_reset", rc);
    auto cc0 = std::chrono::steady_clock::now();
    while (std::fabs(energyDiff) > o.tolerance) {
        if (iterations > max_iters) {
            printf("\n\n\nMax iterations reached. Halting.\n\n\n");
            break;
        }
        if ((rc = ccd_iterate(ctx, mixing, &corrCCD))) return die("ccd_iterate", rc);
        energyDiff = corrCCD - prev;
        prev = corrCCD;
        if (o.verbose >= 1) {
            printf("iterations: %lu\n", iterations);
            printf("corrCCD: %.17g\n", corrCCD);
            printf("energy diff: %.17g\n", energyDiff);
            printf("corrCCD/A: %.17g\n", corrCCD / A);
            printf("energy diff/A: %.17g\n", energyDiff / A);
        }
        iterations++;
    }
    auto cc1 = std::chrono::steady_clock::now();
    printf("iterations: %lu\n", iterations);
    printf("CCDCorr: %.17g\n", corrCCD);
    printf("CCDCorr/A: %.17g\n", corrCCD / A);
    printf("E_CCD: %.17g\n", eref + corrCCD);
    printf("E_CCD/A: %.17g\n", (eref + corrCCD) / A);
    printf("Final time: %g s\n", std::chrono::duration<double>(cc1 - t_total).count());
    printf("cc_diagrams_time: %f\n", std::chrono::duration<double>(cc1 - cc0).count());
    ccd_ctx_destroy(ctx);
    ccd_model_destroy(model);
    return 0;
}
