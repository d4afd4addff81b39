// khs_run — command-line driver with the reference's CLIs:
//   khs_run <config.ini>                      ymgpu2d/main_ym.cu:18-24
//   khs_run --emhd [k_mode] [perturb_amp]     emhdgpu2d/main_coupled.cu:227-229
// (an .ini with "model = emhd" selects the EMHD model as well).  It runs the whole simulation through the C ABI:
// config parse, init (with optional eigenmode seed), output directory, energy
// checks and halting every energy_stride steps, CSV snapshots every
// export_stride steps (main_ym.cu:143-176, 244-491).
// Options (no counterpart in the reference): --checkpoint-every N writes the
// binary checkpoint <output dir>/checkpoint.khs every N steps;
// --restart FILE continues a run from such a checkpoint.
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/khs.h"

namespace fs = std::filesystem;

static int fail(const char* what) {
    std::cerr << "ERROR: " << what << ": " << khs_last_error() << "\n";
    return 1;
}

int main(int argc, char** argv) {
    std::string restart;
    int ckpt_every = 0;
    std::vector<char*> pos;
    for (int i = 0; i < argc; ++i) {
        if (std::strcmp(argv[i], "--restart") == 0 && i + 1 < argc) { restart = argv[++i]; continue; }
        if (std::strcmp(argv[i], "--checkpoint-every") == 0 && i + 1 < argc) { ckpt_every = std::atoi(argv[++i]); continue; }
        pos.push_back(argv[i]);
    }
    argc = (int)pos.size();
    argv = pos.data();
    if (argc < 2) {
        std::cerr << "Usage: khs_run <config.ini> [--checkpoint-every N] [--restart FILE]\n"
                  << "       khs_run --emhd [k_mode] [perturb_amp] [--checkpoint-every N] [--restart FILE]\n"
                  << "  Config is a key=value file with the ymgpu2d example.ini keys.\n";
        return 1;
    }
    khs_config cfg;
    if (std::strcmp(argv[1], "--emhd") == 0) {
        // main_coupled.cu:228-229: k_mode = atoi(argv[1]) (default 1),
        // perturb_amp = atof(argv[2]) (default 0.05)
        if (khs_config_defaults(&cfg) != KHS_OK) return fail("config");
        if (khs_config_set(&cfg, "model", "emhd") != KHS_OK) return fail("config");
        const std::string k = (argc > 2) ? std::to_string(std::atoi(argv[2])) : "1";
        if (khs_config_set(&cfg, "k_mode", k.c_str()) != KHS_OK) return fail("config");
        if (argc > 3 && khs_config_set(&cfg, "perturb_amp", argv[3]) != KHS_OK) return fail("config");
        if (khs_derive(&cfg) != KHS_OK) return fail("config");
    } else if (khs_load_ini(argv[1], &cfg) != KHS_OK) {
        return fail("config");
    }
    const bool emhd = cfg.model == KHS_MODEL_EMHD;
    const char* snap_prefix = emhd ? "coupled_" : "ym_";

    static const char* mode_names[] = {"NAB_LINEAR", "NAB_CIRC", "EMHD_KH", "NAB_DTANH",
                                       "NAB_STEP", "NAB_TANH_COSAZ", "NAB_CIRC_AZ2"};
    if (emhd)
        std::cout << "===============================================================\n"
                  << " MI355X coupled solver: 2D EMHD (" << cfg.species << " fluid"
                  << (cfg.species > 1 ? "s" : "") << ")\n"
                  << " k_mode = " << cfg.k_mode << "  perturb_amp = " << cfg.em_perturb_amp
                  << "  NX=" << cfg.nx << "  NZ=" << cfg.nz << "  DT=" << cfg.dt << "\n"
                  << "===============================================================\n";
    else
        std::cout << "================================================================\n"
                  << " MI355X SU(2) Yang-Mills KH solver — config: " << argv[1] << "\n"
                  << " k_mode=" << cfg.k_mode << "  alpha_YM=" << cfg.alpha_YM
                  << "  perturb_amp=" << cfg.perturb_amp << "  V0=" << cfg.V0
                  << "  mode=" << mode_names[cfg.run_mode] << "\n"
                  << " NX=" << cfg.nx << "  NZ=" << cfg.nz << "  DX=" << cfg.dx << "  DZ=" << cfg.dz
                  << "  DT=" << cfg.dt << "  steps=" << cfg.total_steps
                  << "  precision=" << (cfg.precision == KHS_FP32 ? "fp32" : "fp64") << "\n"
                  << "================================================================\n";

    khs_ctx* ctx = nullptr;
    if (khs_create(&cfg, &ctx) != KHS_OK) return fail("create");

    char dirbuf[1024];
    if (khs_output_dir(&cfg, dirbuf, sizeof(dirbuf)) != KHS_OK) return fail("output dir");
    std::string out_dir = dirbuf;
    fs::create_directories(out_dir);
    if (restart.empty())
        for (auto& e : fs::directory_iterator(out_dir)) {
            auto nm = e.path().filename().string();
            if (nm.rfind(snap_prefix, 0) == 0 && e.path().extension() == ".csv") fs::remove(e.path());
        }

    if (khs_init(ctx, nullptr) != KHS_OK) return fail("init");
    if (!restart.empty()) {
        if (khs_checkpoint_read(ctx, restart.c_str()) != KHS_OK) return fail("restart");
        int s0 = 0;
        khs_get_step(ctx, &s0, nullptr);
        std::cout << " Restarted from " << restart << " at step " << s0 << "\n";
    }
    double E0 = 0.0;
    khs_energy0(ctx, &E0);
    if (khs_energy_csv_open(ctx, (out_dir + "/energy.csv").c_str()) != KHS_OK) return fail("energy.csv");
    std::cout << " Initial energy: " << E0 << "\n";
    if (restart.empty() && cfg.export_stride > 0 && khs_write_snapshot(ctx, out_dir.c_str(), 0) != KHS_OK)
        return fail("snapshot");

    const int total = cfg.total_steps;
    const int es = cfg.energy_stride, xs = cfg.export_stride;
    int step = 0, halted = 0;
    khs_get_step(ctx, &step, nullptr);
    const std::string ckpt_path = out_dir + "/checkpoint.khs";
    while (step < total && !halted) {
        int next = total;
        if (es > 0) next = std::min(next, (step / es + 1) * es);
        if (xs > 0) next = std::min(next, (step / xs + 1) * xs);
        if (ckpt_every > 0) next = std::min(next, (step / ckpt_every + 1) * ckpt_every);
        if (khs_step(ctx, next - step, &halted) != KHS_OK) return fail("step");
        khs_get_step(ctx, &step, nullptr);
        if (es > 0 && step % es == 0) {
            double E = 0.0;
            khs_energy(ctx, &E);
            double t = step * cfg.dt;
            std::cout << " Step " << std::setw(7) << step << " | t=" << std::fixed << std::setprecision(1)
                      << t << " | E/E0=" << std::setprecision(4) << E / E0 << "     \r" << std::flush;
            if (halted) {
                if (std::isnan(E) || std::isinf(E)) std::cout << "\n[HALT] NaN/Inf at t=" << t << "\n";
                else std::cout << "\n[HALT] E/E0=" << E / E0 << " > " << cfg.energy_thresh << " at t=" << t << "\n";
            }
        }
        if (!halted && xs > 0 && step % xs == 0 && khs_write_snapshot(ctx, out_dir.c_str(), step) != KHS_OK)
            return fail("snapshot");
        if (!halted && ckpt_every > 0 && step % ckpt_every == 0 &&
            khs_checkpoint_write(ctx, ckpt_path.c_str()) != KHS_OK)
            return fail("checkpoint");
    }
    if (!halted) std::cout << "\n Run completed normally.\n";
    khs_energy_csv_close(ctx);
    khs_destroy(ctx);
    return 0;
}
