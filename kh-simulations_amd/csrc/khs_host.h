// khs_host.h — host-side internals of libkhs (config, init, context).
#pragma once

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/khs.h"

namespace khs {

// Error carried to the C ABI: code + message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ---- config (khs_config.cpp) ----
void config_defaults(khs_config& c);
// returns false for an unknown key; throws Error on an unparsable value
bool config_set(khs_config& c, const std::string& key, const std::string& val);
void config_load(const std::string& path, khs_config& c, std::string* warnings);
void config_derive(khs_config& c);
std::string output_dir_name(const khs_config& c);

// ---- initial state (khs_init.cpp) ----
// Seed profiles (main_ym.cu:178-217), interpolated onto NX in float.
struct SeedProfiles {
    std::vector<double> f[6];   // by, ex, ez, az, qA, qB (empty = absent)
};
SeedProfiles read_seed_file(const std::string& path, int nx);
// Fill the 24 state arrays (reference layout, double) for cfg; kernel_ym_init
// semantics evaluated on the host, then optional white noise.
void host_initial_state(const khs_config& cfg, const SeedProfiles& seed,
                        std::vector<std::vector<double>>& state);
double noise_value(uint64_t seed, int id, uint64_t cell, double amp);
// EMHD initial state (das_kaw equilibrium or the random-phase spectrum).
void host_initial_state_em(const khs_config& cfg, std::vector<std::vector<double>>& state);

}  // namespace khs
