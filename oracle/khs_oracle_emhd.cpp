// khs_oracle_emhd.cpp — TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// FP64 CPU restatement of the emhdgpu2d time step (emhdgpu2d/main_coupled.cu,
// EMHD_Physics.cu, FCT_Sweeps.cu:184-232): one Abelian fluid (n, px, py, pz)
// coupled to full 3-vector E, B, A on a Yee grid, x walled with one-sided
// derivatives at the walls, z periodic, Lorentz sources applied in BOTH
// sweeps.  Extensions (not in the reference, used by the two-fluid periodic
// configuration): periodic x and a second species with charge +1.
// Same conventions as khs_oracle.cpp: reference expression order, selects for
// min/max, -ffp-contract=off; the in-place (racy) initial smoothing of the
// reference is evaluated with Jacobi semantics.
#include <cmath>
#include <cstdint>
#include <vector>
#include <algorithm>

extern "C" {
struct EmParams {
    int nx, nz, periodic_x, nsp;
    double dx, dz, dt, c, eps0;
    double q0, q1;          // species charges (reference: q0 = -1)
};
}

namespace {

// ids: fields 0..8, currents 9..11, species s at 16 + 12 s: n px py pz ux uz sx sy sz
enum { EX = 0, EY, EZ, BX, BY, BZ, AX, AY, AZ, JX, JY, JZ, SP0 = 16, SPN = 12, EDENS = 40, NIDS = 41 };
enum { N_ = 0, PX_, PY_, PZ_, UX_, UZ_, SX_, SY_, SZ_ };

inline double dmin(double a, double b) { return (b < a) ? b : a; }
inline double dmax(double a, double b) { return (a < b) ? b : a; }

struct EmOracle {
    EmParams p;
    size_t N;
    std::vector<double> a[NIDS];
    explicit EmOracle(const EmParams& pp) : p(pp) {
        N = (size_t)p.nx * p.nz;
        for (int i = 0; i < NIDS; ++i) a[i].assign(N, 0.0);
    }
    double* A(int id) { return a[id].data(); }
    double* F(int s, int k) { return a[SP0 + SPN * s + k].data(); }
    size_t I(int x, int z) const { return (size_t)x + (size_t)z * p.nx; }
    double q(int s) const { return s ? p.q1 : p.q0; }
    int xm(int x) const { return p.periodic_x ? (x + p.nx - 1) % p.nx : std::max(0, x - 1); }
    int xp(int x) const { return p.periodic_x ? (x + 1) % p.nx : std::min(p.nx - 1, x + 1); }
};

// FCT 1-D row, clamped ends (same algorithm as khs_oracle.cpp fct_row).
inline void nu_mu(double eps, double& nu, double& mu) {
    double scrh = dmin(1.0 / 6.0, std::fabs(eps));
    scrh = (1.0 / 3.0) * scrh * scrh;
    nu = (1.0 / 6.0) + (1.0 / 3.0) * (eps * eps - scrh * scrh);
    mu = 0.25 - 0.5 * nu;
}
inline double limiter(double flx, double dm1, double d0, double dp1) {
    double fsgn = std::copysign(1.0, d0);
    double term = fsgn * dm1, terp = fsgn * dp1, fa = std::fabs(flx);
    return fsgn * dmax(0.0, std::fmin(term, std::fmin(fa, terp)));   // CUDA min = fmin
}
void fct_row(int n, const double* u, const double* v, const double* s, double* out, double dx, double dt) {
    std::vector<double> U(n + 6), V(n + 6), S(n + 6), LOR(n + 6), LNR(n + 6), FLX(n + 6);
    for (int e = 0; e < n + 6; ++e) {
        int j = std::min(std::max(e - 3, 0), n - 1);
        U[e] = u[j]; V[e] = v[j]; S[e] = s ? s[j] : 0.0;
    }
    const double dtdx = dt / dx;
    auto epsf = [&](int l, int r) { double vi = 0.5 * (U[l] + U[r]); return vi * dtdx; };
    for (int e = 1; e <= n + 4; ++e) {
        double eL = epsf(e - 1, e), eR = epsf(e, e + 1), nuL, muL, nuR, muR;
        nu_mu(eL, nuL, muL); nu_mu(eR, nuR, muR);
        double fL = 0.5 * eL * (V[e - 1] + V[e]), fR = 0.5 * eR * (V[e] + V[e + 1]);
        double dL = nuL * (V[e] - V[e - 1]), dR = nuR * (V[e + 1] - V[e]);
        LOR[e] = V[e] + dt * S[e] - (fR - fL);
        LNR[e] = LOR[e] + (dR - dL);
    }
    for (int e = 1; e <= n + 3; ++e) {
        double nuR, muR;
        nu_mu(epsf(e, e + 1), nuR, muR);
        FLX[e] = muR * (LOR[e + 1] - LOR[e]);
    }
    for (int i = 0; i < n; ++i) {
        int e = i + 3;
        double dm1 = LNR[e] - LNR[e - 1], d0 = LNR[e + 1] - LNR[e], dp1 = LNR[e + 2] - LNR[e + 1];
        double fl = limiter(FLX[e - 1], LNR[e - 1] - LNR[e - 2], dm1, d0);
        double fr = limiter(FLX[e], dm1, d0, dp1);
        out[i] = LNR[e] - (fr - fl);
    }
}

void update_velocities(EmOracle& o, int s) {          // main_coupled.cu:37-46
    for (size_t i = 0; i < o.N; ++i) {
        double safe_n = dmax(o.F(s, N_)[i], 0.05);
        o.F(s, UX_)[i] = o.F(s, PX_)[i] / safe_n;
        o.F(s, UZ_)[i] = o.F(s, PZ_)[i] / safe_n;
    }
}

// FCT::run_x_sweep (FCT_Sweeps.cu:184-197): n, px, py, pz with the sources
// on the three momenta; walled = clamped ends, periodic = 3-cell ghost wrap.
void x_sweep(EmOracle& o, int s) {
    const int nx = o.p.nx, nz = o.p.nz, H = 3;
    std::vector<double> old[4];
    for (int v = 0; v < 4; ++v) old[v] = o.a[SP0 + SPN * s + v];
    const double* src[4] = {nullptr, o.F(s, SX_), o.F(s, SY_), o.F(s, SZ_)};
    std::vector<double> pu(nx + 2 * H), pv(nx + 2 * H), ps(nx + 2 * H), po(nx + 2 * H);
    for (int z = 0; z < nz; ++z) {
        size_t r = (size_t)z * nx;
        for (int v = 0; v < 4; ++v) {
            const double* sr = src[v] ? src[v] + r : nullptr;
            double* dst = o.F(s, v) + r;
            if (!o.p.periodic_x) {
                fct_row(nx, o.F(s, UX_) + r, old[v].data() + r, sr, dst, o.p.dx, o.p.dt);
            } else {
                auto pack = [&](const double* in, double* out) {
                    for (int x = 0; x < nx; ++x) out[x + H] = in[x];
                    for (int x = 0; x < H; ++x) { out[x] = in[nx - H + x]; out[nx + H + x] = in[x]; }
                };
                pack(o.F(s, UX_) + r, pu.data());
                pack(old[v].data() + r, pv.data());
                if (sr) pack(sr, ps.data());
                fct_row(nx + 2 * H, pu.data(), pv.data(), sr ? ps.data() : nullptr, po.data(), o.p.dx, o.p.dt);
                for (int x = 0; x < nx; ++x) dst[x] = po[x + H];
            }
        }
    }
}

// FCT::run_y_sweep (FCT_Sweeps.cu:199-232): transposed columns, clamped, the
// stale z velocity, the same three sources applied again.
void z_sweep(EmOracle& o, int s) {
    const int nx = o.p.nx, nz = o.p.nz;
    std::vector<double> old[4];
    for (int v = 0; v < 4; ++v) old[v] = o.a[SP0 + SPN * s + v];
    const double* src[4] = {nullptr, o.F(s, SX_), o.F(s, SY_), o.F(s, SZ_)};
    std::vector<double> cu(nz), cv(nz), cs(nz), co(nz);
    for (int x = 0; x < nx; ++x) {
        for (int z = 0; z < nz; ++z) cu[z] = o.F(s, UZ_)[o.I(x, z)];
        for (int v = 0; v < 4; ++v) {
            for (int z = 0; z < nz; ++z) cv[z] = old[v][o.I(x, z)];
            if (src[v]) for (int z = 0; z < nz; ++z) cs[z] = src[v][o.I(x, z)];
            fct_row(nz, cu.data(), cv.data(), src[v] ? cs.data() : nullptr, co.data(), o.p.dz, o.p.dt);
            for (int z = 0; z < nz; ++z) o.F(s, v)[o.I(x, z)] = co[z];
        }
    }
}

void enforce_boundaries(EmOracle& o) {                 // main_coupled.cu:48-63
    for (int z = 0; z < o.p.nz; ++z)
        for (int x = 0; x < o.p.nx; ++x) {
            size_t i = o.I(x, z);
            for (int s = 0; s < o.p.nsp; ++s)
                if (o.F(s, N_)[i] < 0.05) o.F(s, N_)[i] = 0.05;
            if (!o.p.periodic_x && (x == 0 || x == o.p.nx - 1)) {
                for (int s = 0; s < o.p.nsp; ++s) o.F(s, PX_)[i] = 0.0;
                o.A(EY)[i] = 0.0; o.A(EZ)[i] = 0.0; o.A(BX)[i] = 0.0;
            }
        }
}

void currents(EmOracle& o) {                           // main_coupled.cu:99-110
    for (int z = 0; z < o.p.nz; ++z)
        for (int x = 0; x < o.p.nx; ++x) {
            size_t i = o.I(x, z), ixp = o.I(o.xp(x), z), izp = o.I(x, (z == o.p.nz - 1) ? 0 : z + 1);
            double jx = 0, jy = 0, jz = 0;
            for (int s = 0; s < o.p.nsp; ++s) {
                double h = o.q(s) * 0.5;
                double a = h * (o.F(s, PX_)[i] + o.F(s, PX_)[ixp]);
                double b = o.q(s) * o.F(s, PY_)[i];
                double c = h * (o.F(s, PZ_)[i] + o.F(s, PZ_)[izp]);
                if (s == 0) { jx = a; jy = b; jz = c; }
                else { jx = jx + a; jy = jy + b; jz = jz + c; }
            }
            o.A(JX)[i] = jx; o.A(JY)[i] = jy; o.A(JZ)[i] = jz;
        }
}

void ampere(EmOracle& o) {                             // EMHD_Physics.cu:21-54
    const int nx = o.p.nx, nz = o.p.nz;
    const double c2 = o.p.c * o.p.c, dt = o.p.dt;
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x) {
            size_t i = o.I(x, z);
            int zm = (z == 0) ? nz - 1 : z - 1;
            double dBy_dz = (o.A(BY)[i] - o.A(BY)[o.I(x, zm)]) / o.p.dz;
            double curlB_x = -dBy_dz;
            double dBx_dz = (o.A(BX)[i] - o.A(BX)[o.I(x, zm)]) / o.p.dz;
            double dBz_dx, dBy_dx;
            if (o.p.periodic_x) {
                int xm = (x + nx - 1) % nx;
                dBz_dx = (o.A(BZ)[i] - o.A(BZ)[o.I(xm, z)]) / o.p.dx;
                dBy_dx = (o.A(BY)[i] - o.A(BY)[o.I(xm, z)]) / o.p.dx;
            } else if (x == 0) {
                dBz_dx = (o.A(BZ)[o.I(1, z)] - o.A(BZ)[i]) / o.p.dx;
                dBy_dx = (o.A(BY)[o.I(1, z)] - o.A(BY)[i]) / o.p.dx;
            } else {
                dBz_dx = (o.A(BZ)[i] - o.A(BZ)[o.I(x - 1, z)]) / o.p.dx;
                dBy_dx = (o.A(BY)[i] - o.A(BY)[o.I(x - 1, z)]) / o.p.dx;
            }
            double curlB_y = dBx_dz - dBz_dx;
            double curlB_z = dBy_dx;
            o.A(EX)[i] += dt * (c2 * curlB_x - o.A(JX)[i] / o.p.eps0);
            o.A(EY)[i] += dt * (c2 * curlB_y - o.A(JY)[i] / o.p.eps0);
            o.A(EZ)[i] += dt * (c2 * curlB_z - o.A(JZ)[i] / o.p.eps0);
        }
}

void potential(EmOracle& o) {                          // EMHD_Physics.cu:56-65
    for (size_t i = 0; i < o.N; ++i) {
        o.A(AX)[i] -= o.p.dt * o.A(EX)[i];
        o.A(AY)[i] -= o.p.dt * o.A(EY)[i];
        o.A(AZ)[i] -= o.p.dt * o.A(EZ)[i];
    }
}

void faraday(EmOracle& o) {                            // EMHD_Physics.cu:67-98
    const int nx = o.p.nx, nz = o.p.nz;
    const double dt = o.p.dt;
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x) {
            size_t i = o.I(x, z);
            int zp = (z == nz - 1) ? 0 : z + 1;
            double dEy_dz = (o.A(EY)[o.I(x, zp)] - o.A(EY)[i]) / o.p.dz;
            double curlE_x = -dEy_dz;
            double dEx_dz = (o.A(EX)[o.I(x, zp)] - o.A(EX)[i]) / o.p.dz;
            double dEz_dx, dEy_dx;
            if (o.p.periodic_x) {
                int xq = (x + 1) % nx;
                dEz_dx = (o.A(EZ)[o.I(xq, z)] - o.A(EZ)[i]) / o.p.dx;
                dEy_dx = (o.A(EY)[o.I(xq, z)] - o.A(EY)[i]) / o.p.dx;
            } else if (x == nx - 1) {
                dEz_dx = (o.A(EZ)[i] - o.A(EZ)[o.I(x - 1, z)]) / o.p.dx;
                dEy_dx = (o.A(EY)[i] - o.A(EY)[o.I(x - 1, z)]) / o.p.dx;
            } else {
                dEz_dx = (o.A(EZ)[o.I(x + 1, z)] - o.A(EZ)[i]) / o.p.dx;
                dEy_dx = (o.A(EY)[o.I(x + 1, z)] - o.A(EY)[i]) / o.p.dx;
            }
            double curlE_y = dEx_dz - dEz_dx;
            double curlE_z = dEy_dx;
            o.A(BX)[i] -= dt * curlE_x;
            o.A(BY)[i] -= dt * curlE_y;
            o.A(BZ)[i] -= dt * curlE_z;
        }
}

void lorentz(EmOracle& o) {                            // main_coupled.cu:112-136
    const int nx = o.p.nx, nz = o.p.nz;
    for (int s = 0; s < o.p.nsp; ++s)
        for (int z = 0; z < nz; ++z)
            for (int x = 0; x < nx; ++x) {
                size_t i = o.I(x, z);
                int xm = o.xm(x), zm = (z == 0) ? nz - 1 : z - 1;
                double Ex_node = 0.5 * (o.A(EX)[i] + o.A(EX)[o.I(xm, z)]);
                double Ey_node = o.A(EY)[i];
                double Ez_node = 0.5 * (o.A(EZ)[i] + o.A(EZ)[o.I(x, zm)]);
                double Bx_node = 0.5 * (o.A(BX)[i] + o.A(BX)[o.I(x, zm)]);
                double Bz_node = 0.5 * (o.A(BZ)[i] + o.A(BZ)[o.I(xm, z)]);
                double By_node = 0.25 * (o.A(BY)[i] + o.A(BY)[o.I(xm, z)] + o.A(BY)[o.I(x, zm)] + o.A(BY)[o.I(xm, zm)]);
                double n_safe = dmax(o.F(s, N_)[i], 0.05);
                double vx = o.F(s, PX_)[i] / n_safe, vy = o.F(s, PY_)[i] / n_safe, vz = o.F(s, PZ_)[i] / n_safe;
                double q = o.q(s);
                o.F(s, SX_)[i] = q * (Ex_node + (vy * Bz_node - vz * By_node));
                o.F(s, SY_)[i] = q * (Ey_node + (vz * Bx_node - vx * Bz_node));
                o.F(s, SZ_)[i] = q * (Ez_node + (vx * By_node - vy * Bx_node));
            }
}

double energy(EmOracle& o) {                           // main_coupled.cu:197-222
    double total = 0.0;
    for (size_t i = 0; i < o.N; ++i) {
        double ke = 0.0;
        for (int s = 0; s < o.p.nsp; ++s) {
            double n_safe = dmax(o.F(s, N_)[i], 1e-10);
            double px = o.F(s, PX_)[i], py = o.F(s, PY_)[i], pz = o.F(s, PZ_)[i];
            double k = 0.5 * (px * px + py * py + pz * pz) / n_safe;
            ke = (s == 0) ? k : ke + k;
        }
        double em = 0.5 * (o.A(EX)[i] * o.A(EX)[i] + o.A(EY)[i] * o.A(EY)[i] + o.A(EZ)[i] * o.A(EZ)[i] +
                           o.A(BX)[i] * o.A(BX)[i] + o.A(BY)[i] * o.A(BY)[i] + o.A(BZ)[i] * o.A(BZ)[i]);
        o.A(EDENS)[i] = ke + em;
        total += o.A(EDENS)[i];
    }
    return total;
}

void step(EmOracle& o) {                               // main_coupled.cu:322-340
    for (int s = 0; s < o.p.nsp; ++s) update_velocities(o, s);
    for (int s = 0; s < o.p.nsp; ++s) x_sweep(o, s);
    for (int s = 0; s < o.p.nsp; ++s) z_sweep(o, s);
    enforce_boundaries(o);
    currents(o);
    ampere(o); potential(o); faraday(o);
    lorentz(o);
}

}  // namespace

extern "C" {

void* emo_create(const EmParams* p) { return new EmOracle(*p); }
void emo_destroy(void* h) { delete static_cast<EmOracle*>(h); }
double* emo_array(void* h, int id) {
    EmOracle* o = static_cast<EmOracle*>(h);
    if (id < 0 || id >= NIDS) return nullptr;
    return o->A(id);
}
void emo_update_velocities(void* h) { EmOracle& o = *static_cast<EmOracle*>(h); for (int s = 0; s < o.p.nsp; ++s) update_velocities(o, s); }
void emo_x_sweep(void* h, int s) { x_sweep(*static_cast<EmOracle*>(h), s); }
void emo_z_sweep(void* h, int s) { z_sweep(*static_cast<EmOracle*>(h), s); }
void emo_enforce_boundaries(void* h) { enforce_boundaries(*static_cast<EmOracle*>(h)); }
void emo_currents(void* h) { currents(*static_cast<EmOracle*>(h)); }
void emo_ampere(void* h) { ampere(*static_cast<EmOracle*>(h)); }
void emo_potential(void* h) { potential(*static_cast<EmOracle*>(h)); }
void emo_faraday(void* h) { faraday(*static_cast<EmOracle*>(h)); }
void emo_lorentz(void* h) { lorentz(*static_cast<EmOracle*>(h)); }
void emo_step(void* h, int n) { EmOracle& o = *static_cast<EmOracle*>(h); for (int i = 0; i < n; ++i) step(o); }
double emo_energy(void* h) { return energy(*static_cast<EmOracle*>(h)); }

// init_das_kaw_equilibrium (main_coupled.cu:65-94), double math, then the
// 5-point smoothing of By and Ex (:22-35) with Jacobi semantics.
void emo_init_daskaw(void* h, int k_mode, double perturb_amp, int smooth) {
    EmOracle& o = *static_cast<EmOracle*>(h);
    const int nx = o.p.nx, nz = o.p.nz;
    const double dx = o.p.dx, dz = o.p.dz;
    for (int i = 0; i < NIDS; ++i) std::fill(o.a[i].begin(), o.a[i].end(), 0.0);
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x) {
            size_t idx = o.I(x, z);
            double V0 = 0.1, epsilon = 10.0 * dx;
            double k_z = (k_mode * 2.0 * M_PI) / (nz * dz);
            double x_center = (nx * dx) / 2.0;
            double x_node = x * dx;
            double x_norm = (x_node - x_center) / epsilon;
            o.F(0, N_)[idx] = 1.0;
            o.F(0, PZ_)[idx] = V0 * std::tanh(x_norm);
            double x_ex = (x + 0.5) * dx;
            double x_norm_ex = (x_ex - x_center) / epsilon;
            double B0y_ex = -V0 * epsilon * std::log(std::cosh(x_norm_ex));
            double vz_ex = V0 * std::tanh(x_norm_ex);
            o.A(EX)[idx] = vz_ex * B0y_ex;
            double x_by = (x + 0.5) * dx, z_by = (z + 0.5) * dz;
            double x_norm_by = (x_by - x_center) / epsilon;
            double B0y_face = -V0 * epsilon * std::log(std::cosh(x_norm_by));
            o.A(BY)[idx] = B0y_face + perturb_amp * V0 * (1.0 / std::cosh(x_norm_by)) * std::cos(k_z * z_by);
            if (o.p.nsp > 1) o.F(1, N_)[idx] = 1.0;
        }
    if (smooth) {
        for (int id : {BY, EX}) {
            std::vector<double> old = o.a[id];
            for (int z = 0; z < nz; ++z)
                for (int x = 0; x < nx; ++x) {
                    int xm = std::max(0, x - 1), xp = std::min(nx - 1, x + 1);
                    int zm = (z == 0) ? nz - 1 : z - 1, zp = (z == nz - 1) ? 0 : z + 1;
                    o.a[id][o.I(x, z)] = 0.5 * old[o.I(x, z)] +
                        0.125 * (old[o.I(xp, z)] + old[o.I(xm, z)] + old[o.I(x, zp)] + old[o.I(x, zm)]);
                }
        }
    }
}

}  // extern "C"
