// khs_init.cpp — initial state, evaluated once on the host.
//
// kernel_ym_init (ymgpu2d/YM_Init.cu:32-257) evaluates the equilibrium and the
// seed in float math (coshf, tanhf, sinf, cosf, expf, logf, sqrtf of float
// arguments) into double arrays.  The same expressions run here on the host's
// libm, so the state the GPU starts from is a deterministic function of the
// config, independent of the device's float intrinsics.  The seed-profile
// reader and its float interpolation follow main_ym.cu:178-217.
#include <cmath>
#include <cstring>
#include <fstream>

#include "khs_host.h"
#include "khs_layout.h"

namespace khs {

SeedProfiles read_seed_file(const std::string& path, int NX) {
    std::ifstream sf(path, std::ios::binary);
    if (!sf) throw Error(KHS_ERR_IO, "cannot open seed profile: " + path);
    int32_t n_fields_file = 0, nx_file = 0;
    sf.read(reinterpret_cast<char*>(&n_fields_file), sizeof(int32_t));
    sf.read(reinterpret_cast<char*>(&nx_file), sizeof(int32_t));
    if (!sf || n_fields_file < 1 || nx_file < 2)
        throw Error(KHS_ERR_IO, "malformed seed profile header: " + path);
    SeedProfiles sp;
    int n_load = std::min((int)n_fields_file, 6);
    for (int f = 0; f < n_load; ++f) {
        std::vector<float> h(nx_file);
        sf.read(reinterpret_cast<char*>(h.data()), (std::streamsize)nx_file * sizeof(float));
        if (!sf) throw Error(KHS_ERR_IO, "truncated seed profile: " + path);
        std::vector<double> out(NX);
        for (int i = 0; i < NX; ++i) {
            float t = (float)i * (float)(nx_file - 1) / (float)(NX - 1);
            int j = std::min((int)t, nx_file - 2);
            float frc = t - (float)j;
            out[i] = h[j] * (1.0f - frc) + h[j + 1] * frc;
        }
        sp.f[f] = std::move(out);
    }
    return sp;
}

static double log_cosh_stable(double xi) {     // YM_Init.cu:11-18
    const double SAFE_XI = 80.0;
    if (std::fabs(xi) < SAFE_XI) return (double)logf(coshf((float)xi));
    double ax = std::fabs(xi);
    return ax - 0.6931471805599453 + std::log1p(std::exp(-2.0 * ax));
}

double noise_value(uint64_t seed, int id, uint64_t cell, double amp) {
    // splitmix64 finaliser over a key built from (seed, array id, cell)
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + ((uint64_t)(id + 1) << 40) + cell;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    return amp * (2.0 * u - 1.0);
}

void host_initial_state(const khs_config& cfg, const SeedProfiles& seed,
                        std::vector<std::vector<double>>& st) {
    const int nx = cfg.nx, nz = cfg.nz;
    const size_t N = (size_t)nx * nz;
    st.assign(S_COUNT, std::vector<double>(N, 0.0));
    // kernel arguments as main_ym.cu:223-225 passes them
    const double dx = cfg.dx, dz = cfg.dz;
    const int k_mode = cfg.k_mode, run_mode = cfg.run_mode;
    const double perturb_amp = (double)cfg.perturb_amp;
    const double V0 = (double)cfg.V0;
    const double epsilon = cfg.eps;
    const double alpha_YM = (double)(float)cfg.alpha;
    const int init_by1_eq = cfg.init_by1_eq;
    const double vz_edge_taper = (double)cfg.vz_edge_taper;
    const double* s_by = seed.f[0].empty() ? nullptr : seed.f[0].data();
    const double* s_az = seed.f[3].empty() ? nullptr : seed.f[3].data();
    const double* s_qA = seed.f[4].empty() ? nullptr : seed.f[4].data();
    const double* s_qB = seed.f[5].empty() ? nullptr : seed.f[5].data();

    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x) {
            const size_t idx = (size_t)x + (size_t)z * nx;
            const double k_z = (double)((float)k_mode * 2.0f) * (double)M_PI / (nz * dz);
            const double x_c = (nx * dx) * 0.5;
            const double xi = (x * dx - x_c) / epsilon;
            const double z_val = z * dz;
            const float kzz = (float)(k_z * z_val);
            const double sech_xi = (double)(1.0f / coshf((float)xi));
            const double seed_win = (fabsf((float)xi) < 3.0) ? 1.0 : 0.0;
            const double base_seed = perturb_amp * V0 * sech_xi * seed_win;
            const double az1_eq = -V0 * log_cosh_stable(xi);
            double by1 = 0.0, by2 = 0.0, by3 = 0.0, az1 = 0.0, az2 = 0.0, az3 = 0.0;
            double vz_A = +V0 * (double)tanhf((float)xi);
            double qA_amp = 0.0, qB_amp = 0.0;

            if (run_mode == 0) {                          // NAB_LINEAR
                by2 = base_seed * (double)sinf(kzz);
                az1 = az1_eq;
            } else if (run_mode == 1) {                   // NAB_CIRC
                by2 = base_seed * (double)cosf(kzz);
                by3 = base_seed * (double)sinf(kzz);
                az1 = az1_eq;
            } else if (run_mode == 6) {                   // NAB_CIRC_AZ2
                az1 = az1_eq;
                double az_amp;
                if (s_az) {
                    az_amp = perturb_amp * V0 * s_az[x];
                    if (s_by) by2 = perturb_amp * V0 * s_by[x] * (double)cosf(kzz);
                    if (s_by) by3 = perturb_amp * V0 * s_by[x] * (double)sinf(kzz);
                    if (s_qA) qA_amp = perturb_amp * V0 * s_qA[x];
                    if (s_qB) qB_amp = perturb_amp * V0 * s_qB[x];
                } else {
                    double xi_char = (double)(1.0f / sqrtf((float)(alpha_YM * k_z * V0 + (double)1e-12f)));
                    az_amp = perturb_amp * V0 * (double)expf((float)(-0.5 * xi * xi / (xi_char * xi_char)));
                }
                az2 = az_amp * (double)cosf(kzz);
                az3 = az_amp * (double)sinf(kzz);
                if (init_by1_eq) {
                    double mean_lc = (nx * dx) / (4.0 * epsilon) - (double)0.693147f;
                    by1 = -2.0 * V0 * epsilon * ((double)logf(coshf((float)xi)) - mean_lc);
                }
            } else if (run_mode == 2) {                   // EMHD_KH
                by1 = base_seed * (double)sinf(kzz);
                az1 = 0.0;
            } else if (run_mode == 4) {                   // NAB_STEP
                const double Lx = nx * dx, x_phys = x * dx;
                const double s = (x_phys < Lx * (double)(1.0f / 3.0f) ||
                                  x_phys >= Lx * (double)(2.0f / 3.0f)) ? 1.0 : -1.0;
                vz_A = V0 * s;
                const double kx_az1 = 6.0 * (double)M_PI / Lx;
                az1 = -V0 * (double)cosf((float)(kx_az1 * x_phys));
                by2 = perturb_amp * V0 * (double)sinf(kzz);
            } else if (run_mode == 5) {                   // NAB_TANH_COSAZ
                const double Lx = nx * dx, x_ph = x * dx;
                const double xi5 = (x_ph - 0.5 * Lx) / epsilon;
                vz_A = V0 * (double)tanhf((float)xi5);
                const double kx_cos = 2.0 * (double)M_PI / Lx;
                az1 = -V0 * (double)cosf((float)(kx_cos * x_ph));
                by1 = -V0 * kx_cos * (double)sinf((float)(kx_cos * x_ph));
                const double sech5 = (double)(1.0f / coshf((float)xi5));
                const double bs5 = perturb_amp * V0 * sech5;
                by2 = bs5 * (double)cosf(kzz);
                by3 = bs5 * (double)sinf(kzz);
            } else {                                      // NAB_DTANH (mode 3)
                const double L = nx * dx;
                const double xi1 = (x * dx - L * 0.25) / epsilon;
                const double xi2 = (x * dx - L * 0.75) / epsilon;
                const float shear = tanhf((float)xi1) - tanhf((float)xi2) - 1.0f;
                vz_A = V0 * (double)shear;
                az1 = -V0 * (log_cosh_stable(xi1) - log_cosh_stable(xi2));
                by1 = V0 * (double)(tanhf((float)xi1) - tanhf((float)xi2)) / epsilon;
                const double sech1 = (double)(1.0f / coshf((float)xi1));
                const double sech2 = (double)(1.0f / coshf((float)xi2));
                const double base_dt = perturb_amp * V0 * (sech1 + sech2);
                by2 = base_dt * (double)cosf(kzz);
                by3 = base_dt * (double)sinf(kzz);
            }

            if (vz_edge_taper > 0.0 && (run_mode == 1 || run_mode == 6)) {
                const double dt_tap = 3.0;
                const float axi = fabsf((float)xi);
                const float w = 0.5f * (1.0f - tanhf((float)(((double)axi - vz_edge_taper) / dt_tap)));
                vz_A *= (double)w;
                if (init_by1_eq) {
                    const float axi_h = fabsf((float)(xi + 0.5 * dx / epsilon));
                    const double u = ((double)axi_h - vz_edge_taper) / dt_tap;
                    const double I = (double)logf(coshf(axi_h))
                                     - 0.5 * dt_tap * (u + (double)logf(coshf((float)u)) + (double)0.693147f);
                    const double Lxi = (nx * dx) * 0.5 / epsilon;
                    const double vet = vz_edge_taper;
                    const double mean_I = (0.5 * vet * vet - vet * (double)0.693147f
                                           + (Lxi - vet) * (vet - (double)0.693147f)) / Lxi;
                    by1 = -2.0 * V0 * epsilon * (I - mean_I);
                }
            }

            st[S_BY1][idx] = by1; st[S_BY2][idx] = by2; st[S_BY3][idx] = by3;
            st[S_AZ1][idx] = az1; st[S_AZ2][idx] = az2; st[S_AZ3][idx] = az3;
            const double cz = (double)cosf(kzz), sz = (double)sinf(kzz);
            st[S_NA][idx] = 1.0; st[S_PXA][idx] = 0.0; st[S_PZA][idx] = vz_A;
            st[S_Q1A][idx] = 1.0; st[S_Q2A][idx] = qA_amp * cz; st[S_Q3A][idx] = qA_amp * sz;
            st[S_NB][idx] = 1.0; st[S_PXB][idx] = 0.0; st[S_PZB][idx] = -vz_A;
            st[S_Q1B][idx] = -1.0; st[S_Q2B][idx] = qB_amp * cz; st[S_Q3B][idx] = qB_amp * sz;
        }

    // Extension: uniform white noise on colour-2/3 E and By (zero background,
    // so it is representable in the FP32 form too).
    if (cfg.noise_amp > 0.0) {
        const int ids[6] = {S_EX2, S_EX3, S_EZ2, S_EZ3, S_BY2, S_BY3};
        for (int k = 0; k < 6; ++k)
            for (size_t i = 0; i < N; ++i)
                st[ids[k]][i] += noise_value(cfg.noise_seed, ids[k], i, cfg.noise_amp);
    }
}

// --------------------------------------------------------------------------
// EMHD initial state (slots: fields 0..8 = Ex Ey Ez Bx By Bz Ax Ay Az, fluid
// f at 12 + 6 f: n px py pz).  init_das_kaw_equilibrium
// (emhdgpu2d/main_coupled.cu:65-94, double math) followed by the one-time
// 5-point smoothing of By and Ex (:22-35, :279) evaluated with Jacobi
// semantics (the reference kernel updates in place and is racy).
// --------------------------------------------------------------------------
void host_initial_state_em(const khs_config& cfg, std::vector<std::vector<double>>& st) {
    const int nx = cfg.nx, nz = cfg.nz;
    const size_t N = (size_t)nx * nz;
    st.assign(S_COUNT, std::vector<double>(N, 0.0));
    const double dx = cfg.dx, dz = cfg.dz;
    auto I = [&](int x, int z) { return (size_t)x + (size_t)z * nx; };
    const int EXs = 0, BYs = 4, BXs = 3, BZs = 5, N0 = 12, PZ0 = 15, N1 = 18;
    if (cfg.emhd_init == 0) {
        const int k_mode = cfg.k_mode;
        const double perturb_amp = cfg.em_perturb_amp;
        for (int z = 0; z < nz; ++z)
            for (int x = 0; x < nx; ++x) {
                const size_t idx = I(x, z);
                const double V0 = 0.1, epsilon = 10.0 * dx;
                const double k_z = (k_mode * 2.0 * M_PI) / (nz * dz);
                const double x_center = (nx * dx) / 2.0;
                const double x_node = x * dx;
                const double x_norm = (x_node - x_center) / epsilon;
                st[N0][idx] = 1.0;
                st[PZ0][idx] = V0 * std::tanh(x_norm);
                const double x_ex = (x + 0.5) * dx;
                const double x_norm_ex = (x_ex - x_center) / epsilon;
                const double B0y_ex = -V0 * epsilon * std::log(std::cosh(x_norm_ex));
                const double vz_ex = V0 * std::tanh(x_norm_ex);
                st[EXs][idx] = vz_ex * B0y_ex;
                const double x_by = (x + 0.5) * dx, z_by = (z + 0.5) * dz;
                const double x_norm_by = (x_by - x_center) / epsilon;
                const double B0y_face = -V0 * epsilon * std::log(std::cosh(x_norm_by));
                st[BYs][idx] = B0y_face + perturb_amp * V0 * (1.0 / std::cosh(x_norm_by)) * std::cos(k_z * z_by);
                if (cfg.species > 1) st[N1][idx] = 1.0;
            }
    } else {
        // Extension: quiescent fluids and a random-phase spectrum in By,
        // By = amp sum_{kx,kz=1..K} cos(2 pi (kx x/nx + kz z/nz) + phi(kx,kz)),
        // phases from the counter-based generator keyed by spectrum_seed.
        const int K = cfg.spectrum_kmax;
        const double amp = cfg.spectrum_amp;
        for (size_t i = 0; i < N; ++i) { st[N0][i] = 1.0; if (cfg.species > 1) st[N1][i] = 1.0; }
        std::vector<double> gr((size_t)K * nx), gi((size_t)K * nx);
        for (int kz = 1; kz <= K; ++kz)
            for (int x = 0; x < nx; ++x) {
                double re = 0.0, im = 0.0;
                for (int kx = 1; kx <= K; ++kx) {
                    double ph = 2.0 * M_PI * (noise_value(cfg.spectrum_seed, 7, (uint64_t)(kx * 4096 + kz), 0.5) + 0.5);
                    double th = 2.0 * M_PI * (double)((long)kx * x % nx) / nx + ph;
                    re += std::cos(th);
                    im += std::sin(th);
                }
                gr[(size_t)(kz - 1) * nx + x] = re;
                gi[(size_t)(kz - 1) * nx + x] = im;
            }
        for (int z = 0; z < nz; ++z)
            for (int x = 0; x < nx; ++x) {
                double acc = 0.0;
                for (int kz = 1; kz <= K; ++kz) {
                    double tz = 2.0 * M_PI * (double)((long)kz * z % nz) / nz;
                    acc += gr[(size_t)(kz - 1) * nx + x] * std::cos(tz) - gi[(size_t)(kz - 1) * nx + x] * std::sin(tz);
                }
                st[BYs][I(x, z)] = amp * acc;
            }
    }
    (void)BXs; (void)BZs;
    if (cfg.smooth_init) {
        for (int id : {BYs, EXs}) {
            std::vector<double> old = st[id];
            for (int z = 0; z < nz; ++z)
                for (int x = 0; x < nx; ++x) {
                    const int xm = std::max(0, x - 1), xp = std::min(nx - 1, x + 1);
                    const int zm = (z == 0) ? nz - 1 : z - 1, zp = (z == nz - 1) ? 0 : z + 1;
                    st[id][I(x, z)] = 0.5 * old[I(x, z)] +
                        0.125 * (old[I(xp, z)] + old[I(xm, z)] + old[I(x, zp)] + old[I(x, zm)]);
                }
        }
    }
}

}  // namespace khs
