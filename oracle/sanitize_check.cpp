// sanitize_check.cpp — TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// Drives the FP64 CPU oracle (khs_oracle.cpp, khs_oracle_emhd.cpp) and the
// library's host code behind the C ABI of include/khs.h (khs_cpu_abi.cpp:
// .ini keys, configuration, initial states of every run mode, seed-file
// resampling, energy / Gauss / mode amplitude, energy.csv, the growth-rate
// fitters, error paths) through a set of small runs, for builds with
// AddressSanitizer + UndefinedBehaviorSanitizer and with ThreadSanitizer
// (oracle/Makefile: `make sanitize`).  With threads > 1 every thread runs the
// whole set on its own contexts at once: the C ABI promises that separate
// contexts are independent and that the last-error message is per thread.
//
//   sanitize_check SEED_FILE TMP_DIR THREADS
//
// Prints "sanitize_check ok" and exits 0, or names the failed expectation and
// exits 1; the sanitizers report (and, with halt_on_error, abort) on their own.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../include/khs.h"
#include "khs_oracle.h"

extern "C" {
struct EmParams {
    int nx, nz, periodic_x, nsp;
    double dx, dz, dt, c, eps0;
    double q0, q1;
};
void* emo_create(const EmParams* p);
void emo_destroy(void* h);
void emo_step(void* h, int n);
double emo_energy(void* h);
void emo_init_daskaw(void* h, int k_mode, double perturb_amp, int smooth);
}

namespace {

struct Check {
    int tid;
    int failures = 0;
    void expect(bool ok, const char* what) {
        if (!ok) {
            std::fprintf(stderr, "thread %d: expectation failed: %s\n", tid, what);
            ++failures;
        }
    }
};

struct Keys {
    std::vector<std::pair<const char*, std::string>> kv;
};

// One run through the C ABI: configure, init, step, diagnostics, files.
void run_case(Check& ck, const Keys& keys, const char* seed, const std::string& csv, int steps) {
    khs_config cfg;
    ck.expect(khs_config_defaults(&cfg) == KHS_OK, "config defaults");
    for (auto& kv : keys.kv) {
        if (khs_config_set(&cfg, kv.first, kv.second.c_str()) != KHS_OK) {
            std::fprintf(stderr, "thread %d: key %s = %s: %s\n", ck.tid, kv.first, kv.second.c_str(),
                         khs_last_error());
            ++ck.failures;
        }
    }
    ck.expect(khs_derive(&cfg) == KHS_OK, "derive");
    char dir[256];
    ck.expect(khs_output_dir(&cfg, dir, sizeof dir) == KHS_OK, "output dir name");
    khs_ctx* ctx = nullptr;
    if (khs_create(&cfg, &ctx) != KHS_OK) {
        std::fprintf(stderr, "thread %d: create: %s\n", ck.tid, khs_last_error());
        ++ck.failures;
        return;
    }
    ck.expect(khs_init(ctx, seed) == KHS_OK, "init");
    ck.expect(khs_energy_csv_open(ctx, csv.c_str()) == KHS_OK, "energy csv open");
    const size_t n = (size_t)cfg.nx * (size_t)cfg.nz;
    std::vector<double> a(n);
    ck.expect(khs_get_array(ctx, khs_array_id("By2"), a.data(), n) == KHS_OK, "get By2");
    for (double& v : a) v *= 1.0 + 1e-3;
    ck.expect(khs_set_array(ctx, khs_array_id("By2"), a.data(), n) == KHS_OK, "set By2");
    ck.expect(khs_get_array(ctx, khs_array_id("By2"), a.data(), n - 1) != KHS_OK, "short buffer refused");
    int halted = -1;
    ck.expect(khs_step(ctx, steps, &halted) == KHS_OK, "step");
    ck.expect(halted == 0, "no halt");
    double E = 0, E0 = 0, g[7], amp[2];
    ck.expect(khs_energy(ctx, &E) == KHS_OK && khs_energy0(ctx, &E0) == KHS_OK, "energy");
    ck.expect(std::isfinite(E) && E > 0 && E0 > 0, "finite energy");
    ck.expect(khs_gauss(ctx, g) == KHS_OK, "gauss");
    for (double v : g) ck.expect(std::isfinite(v), "finite Gauss sums");
    ck.expect(khs_mode_amplitude(ctx, 1, amp) == KHS_OK && std::isfinite(amp[0]), "mode amplitude");
    int step = 0;
    double t = 0;
    ck.expect(khs_get_step(ctx, &step, &t) == KHS_OK && step == steps, "step counter");
    ck.expect(khs_energy_csv_row(ctx, step, E) == KHS_OK, "energy csv row");
    ck.expect(khs_energy_csv_close(ctx) == KHS_OK, "energy csv close");
    ck.expect(khs_destroy(ctx) == KHS_OK, "destroy");
}

Keys base(int nx, int nz, int run_mode) {
    Keys k;
    k.kv = {{"nx_override", std::to_string(nx)}, {"nz_override", std::to_string(nz)},
            {"lx_override", std::to_string(0.1 * nx)}, {"lz_override", std::to_string(0.1 * nz)},
            {"alpha_YM", "1"}, {"V0", "0.05"}, {"eps_override", "1.0"}, {"courant_override", "0.4"},
            {"k_mode", "1"}, {"run_mode", std::to_string(run_mode)}, {"perturb_amp", "0.01"},
            {"noise_amp", "1e-6"}, {"energy_stride", "3"}, {"export_stride", "-1"}};
    return k;
}

void scenarios(Check& ck, const char* seed_file, const std::string& tmp) {
    const std::string tag = tmp + "/sanitize_t" + std::to_string(ck.tid);
    // error paths: codes and a per-thread message, never exit()
    {
        khs_config cfg;
        khs_config_defaults(&cfg);
        ck.expect(khs_config_set(&cfg, "no_such_key", "1") != KHS_OK, "unknown key refused");
        ck.expect(std::strstr(khs_last_error(), "no_such_key") != nullptr, "per-thread error message");
        ck.expect(khs_config_set(&cfg, "nx_override", "abc") != KHS_OK, "malformed value refused");
        ck.expect(khs_load_ini((tag + "_missing.ini").c_str(), &cfg) != KHS_OK, "missing ini refused");
        khs_config_set(&cfg, "nx_override", "2");
        ck.expect(khs_derive(&cfg) != KHS_OK, "tiny grid refused");
        khs_config_defaults(&cfg);
        khs_config_set(&cfg, "run_mode", "6");
        khs_config_set(&cfg, "nx_override", "64");
        khs_config_set(&cfg, "nz_override", "32");
        khs_ctx* ctx = nullptr;
        ck.expect(khs_create(&cfg, &ctx) == KHS_OK, "create for the seed test");
        if (ctx) {
            ck.expect(khs_init(ctx, (tag + "_missing.bin").c_str()) != KHS_OK, "missing seed file refused");
            khs_destroy(ctx);
        }
    }
    // an .ini file through the reference's parser
    {
        const std::string ini = tag + ".ini";
        FILE* f = std::fopen(ini.c_str(), "w");
        std::fputs("# example\nnx_override = 48  # inline comment\nnz_override = 32\nrun_mode = 1\n"
                   "V0 = 0.05\nunknown_key = 3\nn_steps = 4\n", f);
        std::fclose(f);
        khs_config cfg;
        khs_config_defaults(&cfg);
        ck.expect(khs_load_ini(ini.c_str(), &cfg) == KHS_OK, "ini load");
        ck.expect(khs_derive(&cfg) == KHS_OK && cfg.nx == 48 && cfg.total_steps == 4, "ini values");
        std::remove(ini.c_str());
    }
    // the filter kinds and walls: keep projection + z-mean + xi_cut (walls)
    {
        Keys k = base(64, 32, 0);
        k.kv.push_back({"kz_keep_max", "4"});
        k.kv.push_back({"suppress_kz0", "1"});
        k.kv.push_back({"xi_cut", "2.5"});
        k.kv.push_back({"freeze_override", "1"});
        run_case(ck, k, nullptr, tag + "_a.csv", 4);
    }
    // bands, hyperdiffusion, sponge, warm terms, evolving Az1 (periodic)
    {
        Keys k = base(48, 64, 6);
        k.kv.push_back({"kz_suppress_max", "2"});
        k.kv.push_back({"kz_suppress_hi", "9"});
        k.kv.push_back({"hyp_diff", "1e-3"});
        k.kv.push_back({"xi_sponge", "1.5"});
        k.kv.push_back({"sigma_sponge", "2.0"});
        k.kv.push_back({"warm_T", "0.01"});
        k.kv.push_back({"freeze_override", "0"});
        run_case(ck, k, nullptr, tag + "_b.csv", 4);
    }
    // every other run mode's initial state, odd shapes
    for (int mode = 1; mode <= 5; ++mode) {
        Keys k = base(37 + mode, 20, mode);
        run_case(ck, k, nullptr, tag + "_m" + std::to_string(mode) + ".csv", 2);
    }
    // mode 6 from the reference's eigenvector seed file, resampled onto 77 points
    {
        Keys k = base(77, 16, 6);
        k.kv.push_back({"vz_edge_taper", "3.0"});
        k.kv.push_back({"init_by1_eq", "1"});
        run_case(ck, k, seed_file, tag + "_s.csv", 2);
    }
    for (const char* suffix : {"_a.csv", "_b.csv", "_m1.csv", "_m2.csv", "_m3.csv", "_m4.csv", "_m5.csv", "_s.csv"})
        std::remove((tag + suffix).c_str());
    // growth-rate fitters on a pure exponential
    {
        std::vector<double> t(200), amp(200);
        for (int i = 0; i < 200; ++i) {
            t[i] = 0.5 * i;
            amp[i] = 1e-6 * std::exp(0.09 * t[i]);
        }
        double m[4], pl[5];
        ck.expect(khs_fit_growth_max_r2(t.data(), amp.data(), 200, NAN, NAN, 0.3, m) == KHS_OK, "max-R2 fit");
        ck.expect(std::fabs(m[0] - 0.09) < 1e-9, "max-R2 rate");
        ck.expect(khs_fit_growth_plateau(t.data(), amp.data(), 200, 5.0, 10.0, 0.1, pl) == KHS_OK, "plateau fit");
        ck.expect(khs_fit_growth_max_r2(t.data(), amp.data(), 3, NAN, NAN, 0.3, m) == KHS_OK && m[0] == 0.0,
                  "short series");
    }
    // the oracles' own interfaces: YM operators and the EMHD step
    {
        OrcParams p{};
        p.nx = 40; p.nz = 24; p.periodic_x = 1; p.freeze_az1 = 0;
        p.kz_keep_max = 3; p.dx = p.dz = 0.1; p.dt = 0.04; p.c = p.eps0 = p.alpha = 1.0;
        p.V0 = 0.05; p.eps = 1.0; p.hyp_diff = 1e-3;
        void* h = orc_create(&p);
        for (int id = 12; id < 24; ++id) {
            double* a = orc_array(h, id);
            for (int i = 0; i < p.nx * p.nz; ++i) a[i] = (id == 12 || id == 18) ? 1.0 + 0.01 * (i % 7) : 0.001 * (i % 5);
        }
        orc_sources(h);
        orc_step(h, 3);
        double g[7];
        orc_gauss(h, g);
        ck.expect(std::isfinite(orc_energy(h)) && std::isfinite(g[0]), "oracle step");
        orc_destroy(h);
        EmParams e{32, 32, 1, 2, 1.0, 1.0, 0.002, 1.0, 1.0, -1.0, 1.0};
        void* m = emo_create(&e);
        emo_init_daskaw(m, 2, 0.05, 1);
        emo_step(m, 3);
        ck.expect(std::isfinite(emo_energy(m)), "EMHD oracle step");
        emo_destroy(m);
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s SEED_FILE TMP_DIR THREADS\n", argv[0]);
        return 2;
    }
    const int nt = std::max(1, std::atoi(argv[3]));
    std::vector<Check> checks(nt);
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) {
        checks[i].tid = i;
        th.emplace_back([&, i] { scenarios(checks[i], argv[1], argv[2]); });
    }
    for (auto& t : th) t.join();
    int bad = 0;
    for (auto& c : checks) bad += c.failures;
    if (bad) {
        std::fprintf(stderr, "sanitize_check: %d failed expectations\n", bad);
        return 1;
    }
    std::printf("sanitize_check ok (%d threads)\n", nt);
    return 0;
}
