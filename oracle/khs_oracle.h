// khs_oracle.h — TEST INFRASTRUCTURE, NOT PRODUCT CODE.
// C interface of the FP64 CPU oracle (khs_oracle.cpp): parameters of one run
// and the operators of the ymgpu2d step on its state arrays (ids as in
// include/khs.h: 0..23 state, 32.. derived).
#ifndef KHS_ORACLE_H
#define KHS_ORACLE_H

#include <cstdint>

extern "C" {

struct OrcParams {
    int nx, nz;
    int periodic_x, freeze_az1;
    int suppress_kz0, kz_suppress_max, kz_suppress_hi, kz_keep_max;
    double dx, dz, dt, c, eps0, alpha, V0, eps;
    double xi_sponge, sigma_sponge, xi_cut, warm_T, hyp_diff;
};

void* orc_create(const OrcParams* p);
void orc_destroy(void* h);
double* orc_array(void* h, int id);
void orc_sources(void* h);
void orc_step(void* h, int n);
double orc_energy(void* h);
void orc_gauss(void* h, double* out);

}  // extern "C"

#endif
