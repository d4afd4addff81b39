// khs_cpu_abi.cpp — TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// The FP64 CPU oracle behind the library's C ABI (include/khs.h): libkhs_cpu.so
// exports every entry point the GPU library exports, so the same caller (the
// ctypes binding, the timed CPU baseline) can drive either.  Configuration,
// derivation and the initial state use the library's host code (YM_Config.cu
// and kernel_ym_init semantics, khs_config.cpp / khs_init.cpp), the
// growth-rate fits its khs_analysis.cpp; the time step, energy and Gauss
// residual are the oracle's (khs_oracle.cpp, OpenMP).  The YM model only;
// the operator-by-operator entry points, batches, checkpoints, kernel timing
// and the device-pointer energy return KHS_ERR_STATE ("not in the CPU
// backend"), and an EMHD configuration KHS_ERR_CONFIG.
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "../kh-simulations_amd/csrc/khs_host.h"
#include "khs_oracle.h"

namespace {

thread_local std::string g_err;

struct CpuCtx {
    khs_config cfg{};
    void* orc = nullptr;
    long step = 0;
    double E0 = 0.0;
    double last_E = 0.0;           // the last energy check of khs_step and its step
    long last_step = -1;
    bool sources_stale = true;     // the oracle's sources follow the state it steps
    std::ofstream ecsv;
    ~CpuCtx() {
        if (orc) orc_destroy(orc);
    }
    size_t N() const { return (size_t)cfg.nx * cfg.nz; }
};

template <typename F>
int guard(F&& f) {
    try {
        f();
        g_err.clear();
        return KHS_OK;
    } catch (const khs::Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return KHS_ERR_ARG;
    }
}

int not_here(const char* what) {
    g_err = std::string(what) + ": not in the CPU backend";
    return KHS_ERR_STATE;
}

OrcParams orc_params(const khs_config& c) {
    OrcParams p{};
    p.nx = c.nx; p.nz = c.nz;
    p.periodic_x = c.periodic_x; p.freeze_az1 = c.freeze_az1;
    p.suppress_kz0 = c.suppress_kz0; p.kz_suppress_max = c.kz_suppress_max;
    p.kz_suppress_hi = c.kz_suppress_hi; p.kz_keep_max = c.kz_keep_max;
    p.dx = c.dx; p.dz = c.dz; p.dt = c.dt; p.c = c.c; p.eps0 = c.eps_0;
    p.alpha = c.alpha; p.V0 = c.v0; p.eps = c.eps;
    p.xi_sponge = (double)c.xi_sponge; p.sigma_sponge = (double)c.sigma_sponge; p.xi_cut = (double)c.xi_cut;
    p.warm_T = (double)c.warm_T; p.hyp_diff = (double)c.hyp_diff;
    return p;
}

double energy_of(CpuCtx& c) { return orc_energy(c.orc); }

void write_row(CpuCtx& c, long step, double E) {
    if (!c.ecsv.is_open()) return;
    c.ecsv << step << ',' << step * c.cfg.dt << ',' << E << ',' << E / c.E0 << '\n';
    c.ecsv.flush();
}

}  // namespace

struct khs_ctx {
    std::unique_ptr<CpuCtx> impl;
};
struct khs_batch {
    int unused;
};

#define CTX(ctx)                                                        \
    if (!(ctx) || !(ctx)->impl) { g_err = "null context"; return KHS_ERR_ARG; } \
    CpuCtx& c = *(ctx)->impl;

extern "C" {

const char* khs_last_error(void) { return g_err.c_str(); }
const char* khs_version(void) { return "khs 0.1 (CPU oracle backend, FP64 OpenMP)"; }
size_t khs_config_sizeof(void) { return sizeof(khs_config); }

int khs_config_defaults(khs_config* cfg) {
    if (!cfg) { g_err = "null config"; return KHS_ERR_ARG; }
    return guard([&] { khs::config_defaults(*cfg); khs::config_derive(*cfg); });
}
int khs_config_set(khs_config* cfg, const char* key, const char* value) {
    if (!cfg || !key || !value) { g_err = "null argument"; return KHS_ERR_ARG; }
    return guard([&] {
        if (!khs::config_set(*cfg, key, value))
            throw khs::Error(KHS_ERR_CONFIG, std::string("unknown key '") + key + "'");
    });
}
int khs_load_ini(const char* path, khs_config* cfg) {
    if (!path || !cfg) { g_err = "null argument"; return KHS_ERR_ARG; }
    return guard([&] {
        std::string warn;
        khs::config_load(path, *cfg, &warn);
        khs::config_derive(*cfg);
    });
}
int khs_derive(khs_config* cfg) {
    if (!cfg) { g_err = "null config"; return KHS_ERR_ARG; }
    return guard([&] { khs::config_derive(*cfg); });
}
int khs_output_dir(const khs_config* cfg, char* buf, size_t buflen) {
    if (!cfg || !buf || buflen == 0) { g_err = "null argument"; return KHS_ERR_ARG; }
    return guard([&] {
        const std::string d = khs::output_dir_name(*cfg);
        if (d.size() + 1 > buflen) throw khs::Error(KHS_ERR_ARG, "buffer too small");
        std::memcpy(buf, d.c_str(), d.size() + 1);
    });
}

int khs_create(const khs_config* cfg, khs_ctx** out) {
    if (!cfg || !out) { g_err = "null argument"; return KHS_ERR_ARG; }
    *out = nullptr;
    return guard([&] {
        auto c = std::make_unique<CpuCtx>();
        c->cfg = *cfg;
        khs::config_derive(c->cfg);
        if (c->cfg.model != KHS_MODEL_YM) throw khs::Error(KHS_ERR_CONFIG, "CPU backend: YM model only");
        if (c->cfg.precision != KHS_FP64) throw khs::Error(KHS_ERR_CONFIG, "CPU backend: FP64 only");
        const OrcParams p = orc_params(c->cfg);
        c->orc = orc_create(&p);
        auto h = new khs_ctx;
        h->impl = std::move(c);
        *out = h;
    });
}
int khs_destroy(khs_ctx* ctx) {
    delete ctx;
    return KHS_OK;
}

int khs_init(khs_ctx* ctx, const char* seed_path) {
    CTX(ctx);
    return guard([&] {
        const std::string sp = seed_path ? std::string(seed_path) : std::string(c.cfg.seed_profile_file);
        khs::SeedProfiles seed;
        if (!sp.empty()) seed = khs::read_seed_file(sp, c.cfg.nx);
        std::vector<std::vector<double>> host;
        khs::host_initial_state(c.cfg, seed, host);
        for (int i = 0; i < KHS_N_STATE; ++i) std::memcpy(orc_array(c.orc, i), host[i].data(), c.N() * sizeof(double));
        orc_sources(c.orc);
        c.sources_stale = false;
        c.step = 0;
        c.E0 = energy_of(c);
    });
}

int khs_array_id(const char* name) {
    static const char* names[] = {"Ex1", "Ex2", "Ex3", "Ez1", "Ez2", "Ez3", "By1", "By2", "By3",
                                  "Az1", "Az2", "Az3", "nA", "PxA", "PzA", "Q1A", "Q2A", "Q3A",
                                  "nB", "PxB", "PzB", "Q1B", "Q2B", "Q3B"};
    if (!name) return -1;
    for (int i = 0; i < KHS_N_STATE; ++i)
        if (std::strcmp(name, names[i]) == 0) return i;
    return -1;
}

int khs_get_array(khs_ctx* ctx, int id, double* host, size_t n) {
    CTX(ctx);
    if (!host || n != c.N() || id < 0 || id >= KHS_N_STATE) {
        g_err = "khs_get_array: CPU backend holds the 24 state arrays (nx*nz doubles)";
        return KHS_ERR_ARG;
    }
    std::memcpy(host, orc_array(c.orc, id), n * sizeof(double));
    return KHS_OK;
}
int khs_set_array(khs_ctx* ctx, int id, const double* host, size_t n) {
    CTX(ctx);
    if (!host || n != c.N() || id < 0 || id >= KHS_N_STATE) {
        g_err = "khs_set_array: CPU backend holds the 24 state arrays (nx*nz doubles)";
        return KHS_ERR_ARG;
    }
    std::memcpy(orc_array(c.orc, id), host, n * sizeof(double));
    c.sources_stale = true;
    return KHS_OK;
}

// Steps with the reference's energy check every energy_stride steps and its
// halting rule (main_ym.cu:457-475).
int khs_step(khs_ctx* ctx, int nsteps, int* halted) {
    CTX(ctx);
    if (halted) *halted = 0;
    return guard([&] {
        if (c.sources_stale) {
            orc_sources(c.orc);
            c.sources_stale = false;
        }
        const int stride = c.cfg.energy_stride;
        for (int s = 0; s < nsteps; ++s) {
            orc_step(c.orc, 1);
            ++c.step;
            if (stride > 0 && c.step % stride == 0) {
                const double E = energy_of(c);
                c.last_E = E;
                c.last_step = c.step;
                write_row(c, c.step, E);
                if (std::isnan(E) || std::isinf(E) || E / c.E0 > c.cfg.energy_thresh) {
                    if (halted) *halted = 1;
                    break;
                }
            }
        }
    });
}
int khs_get_step(khs_ctx* ctx, int* step, double* time) {
    CTX(ctx);
    if (step) *step = (int)c.step;
    if (time) *time = c.step * c.cfg.dt;
    return KHS_OK;
}
int khs_energy(khs_ctx* ctx, double* E) {
    CTX(ctx);
    if (!E) { g_err = "null argument"; return KHS_ERR_ARG; }
    *E = energy_of(c);
    return KHS_OK;
}
int khs_last_energy(khs_ctx* ctx, double* E, int* step) {
    CTX(ctx);
    if (!E || !step) { g_err = "null argument"; return KHS_ERR_ARG; }
    *E = c.last_E;
    *step = (int)c.last_step;
    return KHS_OK;
}
int khs_energy0(khs_ctx* ctx, double* E0) {
    CTX(ctx);
    if (!E0) { g_err = "null argument"; return KHS_ERR_ARG; }
    *E0 = c.E0;
    return KHS_OK;
}
int khs_gauss(khs_ctx* ctx, double out[7]) {
    CTX(ctx);
    if (!out) { g_err = "null argument"; return KHS_ERR_ARG; }
    orc_gauss(c.orc, out);
    return KHS_OK;
}
int khs_gauss_residual(khs_ctx* ctx, double rms[3], double rel[3]) {
    if (!rms || !rel) { g_err = "null argument"; return KHS_ERR_ARG; }
    double g[7];
    const int rc = khs_gauss(ctx, g);
    if (rc != KHS_OK) return rc;
    for (int a = 0; a < 3; ++a) {
        rms[a] = g[a];
        rel[a] = g[a] / g[6];
    }
    return KHS_OK;
}
// remote_timeseries.py:62-77 on the host: mean_x |sum_z (Az2 +- i Az3) e^{-2 pi i kz z / NZ}| / NZ
int khs_mode_amplitude(khs_ctx* ctx, int kz, double out[2]) {
    CTX(ctx);
    if (!out || kz < 0) { g_err = "bad argument"; return KHS_ERR_ARG; }
    const int nx = c.cfg.nx, nz = c.cfg.nz;
    const double* a2 = orc_array(c.orc, KHS_AZ2);
    const double* a3 = orc_array(c.orc, KHS_AZ3);
    double s0 = 0.0, s1 = 0.0;
    for (int x = 0; x < nx; ++x) {
        double re = 0, im = 0, rc = 0, ic = 0;
        for (int z = 0; z < nz; ++z) {
            const double ph = 2.0 * M_PI * (double)((long)kz * z % nz) / (double)nz;
            const double cs = std::cos(ph), sn = std::sin(ph);
            const double u = a2[x + (size_t)z * nx], v = a3[x + (size_t)z * nx];
            re += u * cs + v * sn;  im += v * cs - u * sn;     // (u + i v) e^{-i ph}
            rc += u * cs - v * sn;  ic += -v * cs - u * sn;    // (u - i v) e^{-i ph}
        }
        s0 += std::sqrt(re * re + im * im);
        s1 += std::sqrt(rc * rc + ic * ic);
    }
    out[0] = s0 / ((double)nx * nz);
    out[1] = s1 / ((double)nx * nz);
    return KHS_OK;
}

int khs_energy_csv_open(khs_ctx* ctx, const char* path) {
    CTX(ctx);
    if (!path) { g_err = "null argument"; return KHS_ERR_ARG; }
    c.ecsv.open(path);
    if (!c.ecsv) { g_err = std::string("cannot write ") + path; return KHS_ERR_IO; }
    c.ecsv << std::scientific;
    c.ecsv.precision(8);
    c.ecsv << "step,time,E_total,E_ratio\n0,0," << c.E0 << ",1.0\n";
    c.ecsv.flush();
    return KHS_OK;
}
int khs_energy_csv_row(khs_ctx* ctx, int step, double E) {
    CTX(ctx);
    write_row(c, step, E);
    return KHS_OK;
}
int khs_energy_csv_close(khs_ctx* ctx) {
    CTX(ctx);
    if (c.ecsv.is_open()) c.ecsv.close();
    return KHS_OK;
}
int khs_sync(khs_ctx* ctx) {
    CTX(ctx);
    (void)c;
    return KHS_OK;
}

// ---- not in the CPU backend ----
int khs_get_array_buf(khs_ctx*, int, int, double*, size_t) { return not_here("khs_get_array_buf"); }
int khs_fct_x_sweep(khs_ctx*, int) { return not_here("khs_fct_x_sweep"); }
int khs_fct_z_sweep(khs_ctx*, int) { return not_here("khs_fct_z_sweep"); }
int khs_maxwell(khs_ctx*) { return not_here("khs_maxwell"); }
int khs_fluid_velocities(khs_ctx*, int, double*, double*, size_t) { return not_here("khs_fluid_velocities"); }
int khs_sources(khs_ctx*, int, double*, double*, double*, size_t) { return not_here("khs_sources"); }
int khs_enforce_boundaries(khs_ctx*) { return not_here("khs_enforce_boundaries"); }
int khs_filters(khs_ctx*) { return not_here("khs_filters"); }
int khs_commit(khs_ctx*) { return not_here("khs_commit"); }
int khs_energy_noalloc(const void* const*, int, int, int, void*, double*) { return not_here("khs_energy_noalloc"); }
int khs_mode_spectrum(khs_ctx*, int, int, double*) { return not_here("khs_mode_spectrum"); }
int khs_write_snapshot(khs_ctx*, const char*, int) { return not_here("khs_write_snapshot"); }
int khs_checkpoint_write(khs_ctx*, const char*) { return not_here("khs_checkpoint_write"); }
int khs_checkpoint_read(khs_ctx*, const char*) { return not_here("khs_checkpoint_read"); }
int khs_timing_enable(khs_ctx*, int) { return not_here("khs_timing_enable"); }
int khs_timing_reset(khs_ctx*) { return not_here("khs_timing_reset"); }
int khs_kernel_times(khs_ctx*, double*, int*, int) { return not_here("khs_kernel_times"); }
const char* khs_kernel_name(int) { return ""; }
int khs_get_tuning(khs_ctx*, int*) { return not_here("khs_get_tuning"); }
int khs_batch_create(const khs_config*, int, int, int, khs_batch** out) {
    if (out) *out = nullptr;
    return not_here("khs_batch_create");
}
int khs_batch_destroy(khs_batch*) { return KHS_OK; }
int khs_batch_init(khs_batch*) { return not_here("khs_batch_init"); }
int khs_batch_run(khs_batch*, int, int*, int*) { return not_here("khs_batch_run"); }
int khs_batch_run_packed(khs_batch*, int, int*, int*) { return not_here("khs_batch_run_packed"); }
int khs_batch_run_eager(khs_batch*, int, int*, int*) { return not_here("khs_batch_run_eager"); }
int khs_batch_run_concurrent(khs_batch*, int, int, int*, int*) { return not_here("khs_batch_run_concurrent"); }
int khs_batch_get_array(khs_batch*, int, int, double*, size_t) { return not_here("khs_batch_get_array"); }
int khs_batch_set_array(khs_batch*, int, int, const double*, size_t) { return not_here("khs_batch_set_array"); }
int khs_batch_energy(khs_batch*, int, double*, double*) { return not_here("khs_batch_energy"); }
int khs_batch_set_step_limit(khs_batch*, int, int) { return not_here("khs_batch_set_step_limit"); }
int khs_batch_mode_amplitude(khs_batch*, int, int, double*) { return not_here("khs_batch_mode_amplitude"); }

}  // extern "C"
