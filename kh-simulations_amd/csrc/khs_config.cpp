// khs_config.cpp — .ini parser and derived run parameters.
//
// Keys, defaults and parse semantics follow ymgpu2d/YM_Config.cuh:6-50 and
// YM_Config.cu:15-136 (numbers through std::stoi / std::stof into int / float
// fields, inline '#' comments stripped, unknown keys warned).  The derived grid
// and cadence follow main_ym.cu:51-83 and :244-249, evaluated in double from
// the float fields exactly as the reference does (so DT at courant 0.4,
// dx 0.1 is 0.0400000011920929 and target_tu = 80 gives 1999 steps).
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>

#include "khs_host.h"

namespace khs {

static std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    size_t b = s.find_last_not_of(" \t\r\n");
    return (a == std::string::npos) ? "" : s.substr(a, b - a + 1);
}

void config_defaults(khs_config& c) {
    std::memset(&c, 0, sizeof(c));
    c.k_mode = 1;
    c.alpha_YM = 0.5f;
    c.perturb_amp = 0.001f;
    c.run_mode = 6;
    c.V0 = 0.1f;
    c.xi_sponge = 0.0f;
    c.sigma_sponge = 5.0f;
    c.xi_cut = 0.0f;
    c.warm_T = 0.0f;
    c.freeze_override = -1;
    c.suppress_kz0 = 0;
    c.hyp_diff = 0.0f;
    c.kz_suppress_max = 0;
    c.eps_override = -1.0f;
    c.kz_suppress_hi = 0;
    c.vz_edge_taper = 0.0f;
    c.init_by1_eq = 0;
    c.nz_override = -1;
    c.nx_override = -1;
    c.courant_override = -1.0f;
    c.lz_override = -1.0f;
    c.lx_override = -1.0f;
    c.target_tu = -1.0f;
    c.run_tag[0] = '\0';
    c.seed_profile_file[0] = '\0';
    c.n_steps = 0;
    c.noise_amp = 0.0;
    c.noise_seed = 0;
    c.kz_keep_max = 0;
    c.precision = KHS_FP64;
    c.energy_stride_override = 0;
    c.export_stride_override = 0;
    c.fused_sweep = -1;
    c.autotune = 1;
    c.x_segment = 0;
    c.z_chunk = 0;
    c.model = KHS_MODEL_YM;
    c.species = 1;
    c.periodic = 0;
    c.charge1 = -1.0;
    c.charge2 = 1.0;
    c.emhd_init = 0;
    c.smooth_init = 1;
    c.spectrum_kmax = 32;
    c.spectrum_amp = 1e-5;
    c.spectrum_seed = 2;
    c.em_perturb_amp = 0.05;
    c.mfma_filter = 1;
    c.z_prefetch = -1;
    c.x_rows = 0;
    c.maxwell_kernel = -1;
    c.x_serial = -1;
    c.xz_variant = -1;
    c.z_groups = -1;
    c.x_lane_cells = -1;
    c.snake_order = 1;
    c.zm_fused = -1;
    c.step_graphs = 1;
    c.step_fusion = -1;
    c.mxz_variant = -1;
    c.packed_shape = -1;
}

static int parse_int(const std::string& k, const std::string& v) {
    try { return std::stoi(v); } catch (...) {
        throw Error(KHS_ERR_CONFIG, "bad integer for key '" + k + "': '" + v + "'");
    }
}
static float parse_float(const std::string& k, const std::string& v) {
    try { return std::stof(v); } catch (...) {
        throw Error(KHS_ERR_CONFIG, "bad number for key '" + k + "': '" + v + "'");
    }
}
static double parse_double(const std::string& k, const std::string& v) {
    try { return std::stod(v); } catch (...) {
        throw Error(KHS_ERR_CONFIG, "bad number for key '" + k + "': '" + v + "'");
    }
}
static void set_str(char* dst, size_t cap, const std::string& k, const std::string& v) {
    if (v.size() + 1 > cap) throw Error(KHS_ERR_CONFIG, "value too long for key '" + k + "'");
    std::memcpy(dst, v.c_str(), v.size() + 1);
}

bool config_set(khs_config& c, const std::string& key, const std::string& val) {
    // reference keys (YM_Config.cu:99-123)
    if      (key == "k_mode")            c.k_mode = parse_int(key, val);
    else if (key == "alpha_YM")          c.alpha_YM = parse_float(key, val);
    else if (key == "perturb_amp") {
        c.perturb_amp = parse_float(key, val);
        c.em_perturb_amp = parse_double(key, val);
    }
    else if (key == "run_mode")          c.run_mode = parse_int(key, val);
    else if (key == "V0")                c.V0 = parse_float(key, val);
    else if (key == "xi_sponge")         c.xi_sponge = parse_float(key, val);
    else if (key == "sigma_sponge")      c.sigma_sponge = parse_float(key, val);
    else if (key == "xi_cut")            c.xi_cut = parse_float(key, val);
    else if (key == "warm_T")            c.warm_T = parse_float(key, val);
    else if (key == "freeze_override")   c.freeze_override = parse_int(key, val);
    else if (key == "suppress_kz0")      c.suppress_kz0 = parse_int(key, val);
    else if (key == "hyp_diff")          c.hyp_diff = parse_float(key, val);
    else if (key == "kz_suppress_max")   c.kz_suppress_max = parse_int(key, val);
    else if (key == "eps_override")      c.eps_override = parse_float(key, val);
    else if (key == "kz_suppress_hi")    c.kz_suppress_hi = parse_int(key, val);
    else if (key == "init_by1_eq")       c.init_by1_eq = parse_int(key, val);
    else if (key == "vz_edge_taper")     c.vz_edge_taper = parse_float(key, val);
    else if (key == "nz_override")       c.nz_override = parse_int(key, val);
    else if (key == "nx_override")       c.nx_override = parse_int(key, val);
    else if (key == "courant_override")  c.courant_override = parse_float(key, val);
    else if (key == "lz_override")       c.lz_override = parse_float(key, val);
    else if (key == "lx_override")       c.lx_override = parse_float(key, val);
    else if (key == "target_tu")         c.target_tu = parse_float(key, val);
    else if (key == "run_tag")           set_str(c.run_tag, sizeof(c.run_tag), key, val);
    else if (key == "seed_profile_file") set_str(c.seed_profile_file, sizeof(c.seed_profile_file), key, val);
    // extension keys
    else if (key == "n_steps")           c.n_steps = parse_int(key, val);
    else if (key == "noise_amp")         c.noise_amp = parse_double(key, val);
    else if (key == "noise_seed")        c.noise_seed = (uint64_t)std::stoull(val);
    else if (key == "kz_keep_max")       c.kz_keep_max = parse_int(key, val);
    else if (key == "precision") {
        if (val == "fp64") c.precision = KHS_FP64;
        else if (val == "fp32") c.precision = KHS_FP32;
        else throw Error(KHS_ERR_CONFIG, "precision must be fp64 or fp32, got '" + val + "'");
    }
    else if (key == "energy_stride")     c.energy_stride_override = parse_int(key, val);
    else if (key == "export_stride")     c.export_stride_override = parse_int(key, val);
    else if (key == "fused_sweep")       c.fused_sweep = parse_int(key, val);
    else if (key == "autotune")          c.autotune = parse_int(key, val);
    else if (key == "x_segment")         c.x_segment = parse_int(key, val);
    else if (key == "model") {
        if (val == "ym") c.model = KHS_MODEL_YM;
        else if (val == "emhd") c.model = KHS_MODEL_EMHD;
        else throw Error(KHS_ERR_CONFIG, "model must be ym or emhd, got '" + val + "'");
    }
    else if (key == "species")           c.species = parse_int(key, val);
    else if (key == "periodic")          c.periodic = parse_int(key, val);
    else if (key == "charge1")           c.charge1 = parse_double(key, val);
    else if (key == "charge2")           c.charge2 = parse_double(key, val);
    else if (key == "emhd_init") {
        if (val == "daskaw" || val == "0") c.emhd_init = 0;
        else if (val == "spectrum" || val == "1") c.emhd_init = 1;
        else throw Error(KHS_ERR_CONFIG, "emhd_init must be daskaw or spectrum");
    }
    else if (key == "z_groups")          c.z_groups = parse_int(key, val);
    else if (key == "xz_variant")        c.xz_variant = parse_int(key, val);
    else if (key == "x_serial")          c.x_serial = parse_int(key, val);
    else if (key == "maxwell_kernel")    c.maxwell_kernel = parse_int(key, val);
    else if (key == "x_rows")            c.x_rows = parse_int(key, val);
    else if (key == "z_prefetch")        c.z_prefetch = parse_int(key, val);
    else if (key == "mfma_filter")       c.mfma_filter = parse_int(key, val);
    else if (key == "x_lane_cells")      c.x_lane_cells = parse_int(key, val);
    else if (key == "snake_order")       c.snake_order = parse_int(key, val);
    else if (key == "zm_fused")          c.zm_fused = parse_int(key, val);
    else if (key == "step_graphs")       c.step_graphs = parse_int(key, val);
    else if (key == "step_fusion")       c.step_fusion = parse_int(key, val);
    else if (key == "mxz_variant")       c.mxz_variant = parse_int(key, val);
    else if (key == "packed_shape")      c.packed_shape = parse_int(key, val);
    else if (key == "smooth_init")       c.smooth_init = parse_int(key, val);
    else if (key == "spectrum_kmax")     c.spectrum_kmax = parse_int(key, val);
    else if (key == "spectrum_amp")      c.spectrum_amp = parse_double(key, val);
    else if (key == "spectrum_seed")     c.spectrum_seed = (uint64_t)std::stoull(val);
    else if (key == "z_chunk")           c.z_chunk = parse_int(key, val);
    else return false;
    return true;
}

void config_load(const std::string& path, khs_config& c, std::string* warnings) {
    std::ifstream f(path);
    if (!f.is_open()) throw Error(KHS_ERR_CONFIG, "cannot open config file '" + path + "'");
    std::string line;
    int lineno = 0;
    std::ostringstream warn;
    while (std::getline(f, line)) {
        ++lineno;
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        auto eq = line.find('=');
        if (eq == std::string::npos) {
            warn << "WARNING: " << path << ":" << lineno << ": no '=' found, skipping: " << line << "\n";
            continue;
        }
        std::string key = trim(line.substr(0, eq));
        std::string val = trim(line.substr(eq + 1));
        auto hash = val.find('#');
        if (hash != std::string::npos) val = trim(val.substr(0, hash));
        if (!config_set(c, key, val))
            warn << "WARNING: " << path << ":" << lineno << ": unknown key '" << key << "'\n";
    }
    if (warnings) *warnings = warn.str();
    config_derive(c);
}

// emhdgpu2d run (main_coupled.cu:228-260): 256x256, dx = dz = 1, dt = 0.002,
// energy check every 500 steps, snapshot every 1000, halt at E/E0 > 5.
// The kernels address each field array with 32-bit byte offsets.
// Kernels address every array through 32-bit byte offsets, and halo lanes
// that must not store use the offset 0xFFFFFFF0, which the buffer's range
// check drops only while the array is no larger than that: NX*NZ*8 bytes
// <= 0xFFFFFFF0, i.e. NX*NZ <= 2^29 - 2 cells.
static void check_grid_bytes(long nx, long nz) {
    if (nx * nz * (long)sizeof(double) > 0xFFFFFFF0L)
        throw Error(KHS_ERR_CONFIG, "grid too large: NX*NZ must be at most 2^29 - 2 cells (FP64 arrays of at most "
                                    "0xFFFFFFF0 bytes)");
}

static void config_derive_emhd(khs_config& c) {
    c.nz = (c.nz_override > 0) ? c.nz_override : 256;
    c.nx = (c.nx_override > 0) ? c.nx_override : 256;
    c.lx = (c.lx_override > 0.f) ? (double)c.lx_override : (double)c.nx;
    c.lz = (c.lz_override > 0.f) ? (double)c.lz_override : (double)c.nz;
    c.dx = c.lx / c.nx;
    c.dz = c.lz / c.nz;
    c.courant = (c.courant_override > 0.f) ? (double)c.courant_override : 0.002;
    c.dt = c.courant * c.dx;
    c.v0 = 0.1;
    c.alpha = 0.0;
    c.eps = 10.0 * c.dx;
    c.c = 1.0;
    c.eps_0 = 1.0;
    c.periodic_x = c.periodic ? 1 : 0;
    c.freeze_az1 = 0;
    c.total_steps = (c.target_tu > 0.0f) ? (int)((double)c.target_tu / c.dt) : 500000;
    if (c.n_steps > 0) c.total_steps = c.n_steps;
    c.energy_stride = (c.energy_stride_override > 0) ? c.energy_stride_override : 500;
    c.export_stride = (c.export_stride_override > 0) ? c.export_stride_override : 1000;
    if (c.export_stride_override < 0) c.export_stride = 0;
    c.energy_thresh = 5.0;
    if (c.nx < 4 || c.nz < 4) throw Error(KHS_ERR_CONFIG, "grid too small");
    check_grid_bytes(c.nx, c.nz);
    if (c.xz_variant > 6) throw Error(KHS_ERR_CONFIG, "xz_variant must be -1..6 for the EMHD model");
    if (c.maxwell_kernel > 3) throw Error(KHS_ERR_CONFIG, "maxwell_kernel must be -1..3 for the EMHD model");
    if (c.species != 1 && c.species != 2) throw Error(KHS_ERR_CONFIG, "species must be 1 or 2");
    if (c.x_segment != 0 && c.x_segment != 64 && c.x_segment != 128 && c.x_segment != 256)
        throw Error(KHS_ERR_CONFIG, "x_segment must be 0, 64, 128 or 256");
    if (c.z_chunk != 0 && c.z_chunk != 16 && c.z_chunk != 32 && c.z_chunk != 64)
        throw Error(KHS_ERR_CONFIG, "z_chunk must be 0, 16, 32 or 64");
    if (c.z_prefetch < -1 || c.z_prefetch > 2) throw Error(KHS_ERR_CONFIG, "z_prefetch must be -1, 0, 1 or 2");
    if (c.x_rows != 0 && c.x_rows != 1 && c.x_rows != 4 && c.x_rows != 16)
        throw Error(KHS_ERR_CONFIG, "x_rows must be 0, 1, 4 or 16");
    if (c.maxwell_kernel < -1 || c.maxwell_kernel > 10) throw Error(KHS_ERR_CONFIG, "maxwell_kernel must be -1..10");
    if (c.x_serial < -1 || c.x_serial > 1) throw Error(KHS_ERR_CONFIG, "x_serial must be -1, 0 or 1");
    if (c.xz_variant < -1 || c.xz_variant > 14) throw Error(KHS_ERR_CONFIG, "xz_variant must be -1..14");
    if (c.xz_variant >= 13 && c.warm_T > 0.0f) throw Error(KHS_ERR_CONFIG, "xz_variant 13, 14 need warm_T = 0");
    if (c.fused_sweep < -1 || c.fused_sweep > 1) throw Error(KHS_ERR_CONFIG, "fused_sweep must be -1, 0 or 1");
    if (c.z_groups < -1 || c.z_groups > 1) throw Error(KHS_ERR_CONFIG, "z_groups must be -1, 0 or 1");
    if (c.x_lane_cells != -1 && c.x_lane_cells != 0 && c.x_lane_cells != 4 && c.x_lane_cells != 8)
        throw Error(KHS_ERR_CONFIG, "x_lane_cells must be -1, 0, 4 or 8");
    if (c.snake_order != 0 && c.snake_order != 1) throw Error(KHS_ERR_CONFIG, "snake_order must be 0 or 1");
    if (c.zm_fused < -1 || c.zm_fused > 3) throw Error(KHS_ERR_CONFIG, "zm_fused must be -1..3");
    if (c.emhd_init == 1 && (c.spectrum_kmax < 1 || 2 * c.spectrum_kmax >= std::min(c.nx, c.nz)))
        throw Error(KHS_ERR_CONFIG, "spectrum_kmax must satisfy 1 <= K < min(nx, nz)/2");
}

void config_derive(khs_config& c) {
    if (c.model == KHS_MODEL_EMHD) { config_derive_emhd(c); return; }
    if (c.model != KHS_MODEL_YM) throw Error(KHS_ERR_CONFIG, "unknown model");
    c.nz = (c.nz_override > 0) ? c.nz_override : 64;
    c.nx = (c.nx_override > 0) ? c.nx_override : 768;
    c.lx = (c.lx_override > 0.f) ? (double)c.lx_override : (double)(6.0 * M_PI);
    c.lz = (c.lz_override > 0.f) ? (double)c.lz_override : (double)(2.0 * M_PI);
    c.dx = c.lx / c.nx;
    c.dz = c.lz / c.nz;
    c.courant = (c.courant_override > 0.f) ? (double)c.courant_override : (double)0.1f;
    c.dt = c.courant * c.dx;
    c.v0 = (double)c.V0;
    c.alpha = (double)c.alpha_YM;
    c.eps = (c.eps_override > 0.f) ? (double)c.eps_override : c.lx / 6.0;
    c.c = 1.0;
    c.eps_0 = 1.0;
    const int m = c.run_mode;
    c.freeze_az1 = (m == 1 || m == 3 || m == 4 || m == 5 || m == 6) ? 1 : 0;
    if (c.freeze_override >= 0) c.freeze_az1 = c.freeze_override;
    c.periodic_x = (m >= 1) ? 1 : 0;
    // main_ym.cu:244-249
    // n_steps (extension) fixes the step count and, like target_tu, selects the
    // physical-time cadences (export every 1 TU, energy check every 0.5 TU).
    const bool timed = (c.target_tu > 0.0f) || (c.n_steps > 0);
    c.total_steps = (c.target_tu > 0.0f) ? (int)((double)c.target_tu / c.dt) : 2000000;
    c.export_stride = timed ? std::max(1, (int)(1.0 / c.dt)) : 20000;
    c.energy_stride = timed ? std::max(1, (int)(0.5 / c.dt)) : 10000;
    c.energy_thresh = (m >= 3 || c.xi_sponge > 0.0) ? 100.0 : 5.0;
    if (c.n_steps > 0) c.total_steps = c.n_steps;
    if (c.energy_stride_override > 0) c.energy_stride = c.energy_stride_override;
    if (c.export_stride_override > 0) c.export_stride = c.export_stride_override;
    if (c.export_stride_override < 0) c.export_stride = 0;

    // constraint checks (YM_Config.cu:119-133) plus the build's own limits
    const int NZ = c.nz, NX = c.nx;
    if (NX < 4 || NZ < 4)
        throw Error(KHS_ERR_CONFIG, "grid too small (NX=" + std::to_string(NX) + ", NZ=" + std::to_string(NZ) + ")");
    check_grid_bytes(NX, NZ);
    if (c.suppress_kz0 && (NZ & (NZ - 1)) != 0)
        throw Error(KHS_ERR_CONFIG, "suppress_kz0=1 requires NZ to be a power of 2 (got NZ=" + std::to_string(NZ) + ")");
    if ((c.kz_suppress_max >= 1 || c.kz_suppress_hi > 0) && (NZ % 32 != 0))
        throw Error(KHS_ERR_CONFIG, "kz_suppress_max/kz_suppress_hi require NZ multiple of 32 (got NZ=" + std::to_string(NZ) + ")");
    if (c.kz_keep_max < 0 || (c.kz_keep_max > 0 && 2 * c.kz_keep_max >= NZ))
        throw Error(KHS_ERR_CONFIG, "kz_keep_max must satisfy 0 <= K < NZ/2");
    if (c.hyp_diff > 0.0f && NZ < 5)
        throw Error(KHS_ERR_CONFIG, "hyp_diff requires NZ >= 5");
    if (c.kz_suppress_hi > NZ / 2 || c.kz_suppress_max > NZ / 2)
        throw Error(KHS_ERR_CONFIG, "kz band edges must not exceed NZ/2");
    if (c.x_segment != 0 && c.x_segment != 64 && c.x_segment != 128 && c.x_segment != 256)
        throw Error(KHS_ERR_CONFIG, "x_segment must be 0, 64, 128 or 256");
    if (c.z_chunk != 0 && c.z_chunk != 16 && c.z_chunk != 32 && c.z_chunk != 64)
        throw Error(KHS_ERR_CONFIG, "z_chunk must be 0, 16, 32 or 64");
    if (c.z_prefetch < -1 || c.z_prefetch > 2) throw Error(KHS_ERR_CONFIG, "z_prefetch must be -1, 0, 1 or 2");
    if (c.x_rows != 0 && c.x_rows != 1 && c.x_rows != 4 && c.x_rows != 16)
        throw Error(KHS_ERR_CONFIG, "x_rows must be 0, 1, 4 or 16");
    if (c.maxwell_kernel < -1 || c.maxwell_kernel > 10) throw Error(KHS_ERR_CONFIG, "maxwell_kernel must be -1..10");
    if (c.x_serial < -1 || c.x_serial > 1) throw Error(KHS_ERR_CONFIG, "x_serial must be -1, 0 or 1");
    if (c.xz_variant < -1 || c.xz_variant > 14) throw Error(KHS_ERR_CONFIG, "xz_variant must be -1..14");
    if (c.xz_variant >= 13 && c.warm_T > 0.0f) throw Error(KHS_ERR_CONFIG, "xz_variant 13, 14 need warm_T = 0");
    if (c.fused_sweep < -1 || c.fused_sweep > 1) throw Error(KHS_ERR_CONFIG, "fused_sweep must be -1, 0 or 1");
    if (c.z_groups < -1 || c.z_groups > 1) throw Error(KHS_ERR_CONFIG, "z_groups must be -1, 0 or 1");
    if (c.x_lane_cells != -1 && c.x_lane_cells != 0 && c.x_lane_cells != 4 && c.x_lane_cells != 8)
        throw Error(KHS_ERR_CONFIG, "x_lane_cells must be -1, 0, 4 or 8");
    if (c.snake_order != 0 && c.snake_order != 1) throw Error(KHS_ERR_CONFIG, "snake_order must be 0 or 1");
    if (c.zm_fused < -1 || c.zm_fused > 3) throw Error(KHS_ERR_CONFIG, "zm_fused must be -1..3");
    if (c.step_fusion < -1 || c.step_fusion > 1) throw Error(KHS_ERR_CONFIG, "step_fusion must be -1, 0 or 1");
    if (c.mxz_variant < -1 || c.mxz_variant > 8) throw Error(KHS_ERR_CONFIG, "mxz_variant must be -1..8");
    if (c.packed_shape < -1 || c.packed_shape > 7) throw Error(KHS_ERR_CONFIG, "packed_shape must be -1..7");
    if (c.run_mode < 0 || c.run_mode > 6)
        throw Error(KHS_ERR_CONFIG, "run_mode must be 0..6");
}

// Output directory name (main_ym.cu:44-49, 143-170; EMHD main_coupled.cu:262-264).
std::string output_dir_name(const khs_config& c) {
    if (c.model == KHS_MODEL_EMHD)
        return "coupled_k" + std::to_string(c.k_mode) + (c.em_perturb_amp == 0.0 ? "_noperturb" : "");
    const int run_mode = c.run_mode;
    const char* mode_tag = (run_mode == 1) ? "_circ"
                         : (run_mode == 2) ? "_emhd"
                         : (run_mode == 3) ? "_dtanh"
                         : (run_mode == 4) ? "_step"
                         : (run_mode == 5) ? "_tanh"
                         : (run_mode == 6) ? "_circ_az2seed" : "";
    const double aYM = c.alpha_YM, p_amp = c.perturb_amp, V0 = c.V0;
    const double xi_sponge = c.xi_sponge, xi_cut = c.xi_cut, hyp_diff = c.hyp_diff;
    std::ostringstream d;
    d << "ym_k" << c.k_mode << "_a" << std::fixed << std::setprecision(3) << aYM;
    if (p_amp == 0.0) d << "_noperturb";
    d << mode_tag;
    if (V0 != 0.1) d << "_v" << std::setprecision(4) << V0;
    if (xi_sponge > 0.0) d << "_sp" << std::setprecision(1) << xi_sponge;
    if (xi_cut > 0.0) d << "_xc" << std::setprecision(1) << xi_cut;
    if (c.freeze_az1 && run_mode == 0) d << "_frz";
    if (c.eps_override > 0.0f) d << "_eps" << std::fixed << std::setprecision(2) << c.eps_override;
    if (c.nx_override > 0) d << "_nx" << c.nx_override;
    if (c.lx_override > 0.f) d << "_lx" << std::fixed << std::setprecision(2) << c.lx_override;
    if (c.suppress_kz0) d << "_nkz0";
    if (c.kz_suppress_max >= 1) d << "_nkz1to" << c.kz_suppress_max;
    if (c.kz_suppress_hi > 0) d << "_bp" << c.kz_suppress_hi;
    if (hyp_diff > 0.0) d << "_hd" << std::scientific << std::setprecision(0) << hyp_diff;
    if (c.run_tag[0]) d << "_" << c.run_tag;
    return d.str();
}

}  // namespace khs
