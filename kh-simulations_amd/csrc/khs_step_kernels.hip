// khs_step_kernels.hip — the per-step hot path on gfx950.
//
//   k_fct_x   directionally split FCT sweep along x (contiguous axis) for all six
//             fluid variables of a species in one launch.  One workgroup owns a
//             256-cell row segment; the segment plus a 3-cell halo is staged in
//             LDS once and every variable reuses the shared interface
//             coefficients (eps, nu, mu depend only on the velocity).  Velocity
//             and the px Lorentz source are evaluated on the fly from the state.
//   k_fct_z   FCT sweep along z (strided axis) without transposes: each lane owns
//             one x column (coalesced along x) and marches down a z chunk,
//             carrying the FCT pipeline (fluxes, LORHOT/LNRHOT, limited fluxes)
//             in registers.  Velocity and the pz / colour-precession sources are
//             evaluated on the fly from the step-start state.  Writes the swept
//             fluid with the density floor and x-wall condition applied.
//   k_maxwell currents + Ampere + potential + Faraday for a 64x16 tile.  The
//             updated E (needed at x+1 and z+1 by Faraday) is produced into LDS
//             for the tile plus a one-cell halo, so E never makes an extra trip
//             through HBM.
//
// Reference semantics: ymgpu2d/FCT_Sweeps.cu:39-177, YM_FCT.cu:30-135,
// YM_Fluid.cu:3-108, YM_Physics.cu:9-151,589-599, main_ym.cu:322-363.
#include "khs_device.h"
#include "khs_launch.h"

namespace khs {

// ==========================================================================
// x sweep
// ==========================================================================
constexpr int XB = 256;            // cells per row segment (= threads)
constexpr int XW = XB + 6;         // segment + 3-cell halo each side

template <typename T>
__global__ __launch_bounds__(XB) void k_fct_x(State<T> S, Stage<T> out, KParams p, int sp0) {
    __shared__ T sU[XW], sS[XW];
    __shared__ T sE[XW], sNU[XW], sMU[XW];
    __shared__ T sV[F_COUNT][XW];      // variable, later the limited interface flux
    __shared__ T sLOR[F_COUNT][XW];
    __shared__ T sLNR[F_COUNT][XW];

    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * XB;
    const int z = blockIdx.y;
    const int sp = sp0 + blockIdx.z;
    const int b = S_NA + sp * F_COUNT;
    const size_t row = (size_t)z * p.nx;
    const T dt = T(p.dt);
    const T dtdx = T(p.dtdx);

    // Phase A: stage variables, velocity and px source for positions
    // x0-3 .. x0+XB+2 (position q <-> extended x = x0-3+q).
    for (int q = tid; q < XW; q += XB) {
        const int xr = map_x(x0 - 3 + q, p);
        const size_t i = row + xr;
        T n = S.a[b + F_N][i];
        T px = S.a[b + F_PX][i];
        T n_safe = kmax(n, T(0.05));
        sU[q] = px / n_safe;                               // kernel_update_velocities
        sS[q] = lorentz_px(S, sp, xr, z, p);
        sV[F_N][q] = n;
        sV[F_PX][q] = px;
#pragma unroll
        for (int v = F_PZ; v < F_COUNT; ++v) sV[v][q] = S.a[b + v][i];
    }
    __syncthreads();

    // Phase B: interface q+1/2 coefficients (compute_eps, compute_nu_mu)
    for (int q = tid; q < XW - 1; q += XB) {
        T vi = T(0.5) * (sU[q] + sU[q + 1]);
        T eps = vi * dtdx;
        T nu, mu;
        fct_nu_mu(eps, nu, mu);
        sE[q] = eps; sNU[q] = nu; sMU[q] = mu;
    }
    __syncthreads();

    // Phase C: transported (LORHOT) and diffused (LNRHOT) low-order values at
    // positions 1 .. XW-2 (FCT_Sweeps.cu:88-103).
    for (int q = 1 + tid; q < XW - 1; q += XB) {
        const T eL = sE[q - 1], eR = sE[q];
        const T nL = sNU[q - 1], nR = sNU[q];
#pragma unroll
        for (int v = 0; v < F_COUNT; ++v) {
            const T vm = sV[v][q - 1], v0 = sV[v][q], vp = sV[v][q + 1];
            T fL = T(0.5) * eL * (vm + v0);
            T fR = T(0.5) * eR * (v0 + vp);
            T dL = nL * (v0 - vm);
            T dR = nR * (vp - v0);
            T src = (v == F_PX) ? sS[q] : T(0.0);
            T lor = v0 + dt * src - (fR - fL);
            sLOR[v][q] = lor;
            sLNR[v][q] = lor + (dR - dL);
        }
    }
    __syncthreads();

    // Phase D: limited antidiffusive flux at interfaces q+1/2, q = 2 .. XW-4
    // (FCT_Sweeps.cu:139, 155-160).  Written over sV (no longer needed).
    for (int q = 2 + tid; q < XW - 3; q += XB) {
        const T mu = sMU[q];
#pragma unroll
        for (int v = 0; v < F_COUNT; ++v) {
            T flx = mu * (sLOR[v][q + 1] - sLOR[v][q]);
            T l0 = sLNR[v][q - 1], l1 = sLNR[v][q], l2 = sLNR[v][q + 1], l3 = sLNR[v][q + 2];
            T fl = fct_limit(flx, l1 - l0, l2 - l1, l3 - l2);
            sV[v][q] = fl;   // safe: Phase D reads only sLOR/sLNR
        }
    }
    __syncthreads();

    // Phase E: final update for the segment cells (positions 3 .. XB+2)
    const int x = x0 + tid;
    if (x < p.nx) {
        const int q = 3 + tid;
#pragma unroll
        for (int v = 0; v < F_COUNT; ++v) {
            T r = sLNR[v][q] - (sV[v][q] - sV[v][q - 1]);
            out.a[sp * F_COUNT + v][row + x] = r;
        }
    }
}

// ==========================================================================
// z sweep (register march)
// ==========================================================================
constexpr int ZB_X = 64;           // columns per workgroup (one wave wide)
constexpr int ZB_C = 4;            // z chunks per workgroup (one per wave)

template <typename T, int TZ>
__global__ __launch_bounds__(ZB_X * ZB_C) void k_fct_z(State<T> S, Stage<T> X, State<T> O,
                                                       KParams p, int sp0) {
    const int x = blockIdx.x * ZB_X + (threadIdx.x % ZB_X);
    const int chunk = blockIdx.y * ZB_C + (threadIdx.x / ZB_X);
    const int z0 = chunk * TZ;
    const int sp = sp0 + blockIdx.z;
    if (x >= p.nx || z0 >= p.nz) return;
    const int nx = p.nx, nz = p.nz;
    const int b = S_NA + sp * F_COUNT;
    const T dt = T(p.dt), dtdz = T(p.dtdz);
    const int zend = min(z0 + TZ, nz);
    const bool wall = !p.periodic_x && (x == 0 || x == nx - 1);

    // pipeline registers (names: *_m = previous row / interface)
    T Vm[F_COUNT], Sm[F_COUNT];      // variable and source at row e-1
    T Fm[F_COUNT], Dm[F_COUNT];      // transport / diffusion flux at e-3/2
    T LORm[F_COUNT];                 // LORHOT(e-2)
    T L3[F_COUNT], L2[F_COUNT], L1[F_COUNT];  // LNRHOT(e-4), (e-3), (e-2)
    T FLXm[F_COUNT];                 // antidiffusive flux at e-5/2
    T FLm[F_COUNT];                  // limited flux at e-7/2
    T Um = T(0), MUm = T(0);         // velocity at e-1, mu at interface e-3/2

    for (int e = z0 - 3; e <= zend + 2; ++e) {
        const int rc = clampi(e, 0, nz - 1);
        const size_t i = (size_t)x + (size_t)rc * nx;
        // ---- load row e: x-swept variables, stale velocity, sources ----
        T V[F_COUNT], Sv[F_COUNT];
#pragma unroll
        for (int v = 0; v < F_COUNT; ++v) V[v] = X.a[sp * F_COUNT + v][i];
        T u;
        {
            T n_safe = kmax(S.a[b + F_N][i], T(0.05));
            u = S.a[b + F_PZ][i] / n_safe;
        }
        Sv[F_N] = T(0); Sv[F_PX] = T(0);
        Sv[F_PZ] = lorentz_pz(S, sp, x, rc, p);
        {
            T q[3];
            precession_at(S, sp, x, rc, p, q);
            Sv[F_Q1] = q[0]; Sv[F_Q2] = q[1]; Sv[F_Q3] = q[2];
        }
        if (e > z0 - 3) {
            // ---- interface (e-1, e) ----
            T vi = T(0.5) * (Um + u);
            T eps = vi * dtdz;
            T nu, mu;
            fct_nu_mu(eps, nu, mu);
            const bool do_lor = (e >= z0 - 1);
            const bool do_flx = (e >= z0);
            const bool do_fl = (e >= z0 + 2);
            const bool do_out = (e >= z0 + 3);
#pragma unroll
            for (int v = 0; v < F_COUNT; ++v) {
                T F = T(0.5) * eps * (Vm[v] + V[v]);
                T D = nu * (V[v] - Vm[v]);
                if (do_lor) {
                    // row e-1: LORHOT / LNRHOT
                    T lor = Vm[v] + dt * Sm[v] - (F - Fm[v]);
                    T lnr = lor + (D - Dm[v]);
                    T flx = T(0);
                    if (do_flx) flx = MUm * (lor - LORm[v]);      // interface e-3/2
                    if (do_fl) {
                        // limited flux at interface e-5/2 (j = e-3)
                        T fl = fct_limit(FLXm[v], L2[v] - L3[v], L1[v] - L2[v], lnr - L1[v]);
                        if (do_out) {
                            T r = L2[v] - (fl - FLm[v]);          // row e-3
                            const int zo = e - 3;
                            size_t io = (size_t)x + (size_t)zo * nx;
                            if (v == F_N) { if (r < T(0.05)) r = T(0.05); }   // YM_Fluid.cu:20-21
                            if (v == F_PX && wall) r = T(0.0);                // YM_Fluid.cu:23-25
                            O.a[b + v][io] = r;
                        }
                        FLm[v] = fl;
                    }
                    FLXm[v] = flx;
                    LORm[v] = lor;
                    L3[v] = L2[v]; L2[v] = L1[v]; L1[v] = lnr;
                }
                Fm[v] = F; Dm[v] = D;
            }
            MUm = mu;
        }
#pragma unroll
        for (int v = 0; v < F_COUNT; ++v) { Vm[v] = V[v]; Sm[v] = Sv[v]; }
        Um = u;
    }
}

// ==========================================================================
// Maxwell: currents + Ampere + potential + Faraday
// ==========================================================================
constexpr int MX = 64, MZ = 16, MZT = 4;   // tile 64 x 16, 64 x 4 threads

template <typename T>
__global__ __launch_bounds__(MX * MZT) void k_maxwell(State<T> S, State<T> O, KParams p) {
    __shared__ T sEx[3][MZ + 1][MX + 1];
    __shared__ T sEz[3][MZ + 1][MX + 1];
    const int nx = p.nx, nz = p.nz;
    const int x0 = blockIdx.x * MX, z0 = blockIdx.y * MZ;
    const int ncol = min(MX, nx - x0), nrow = min(MZ, nz - z0);
    const T dt = T(p.dt), c2 = T(p.c2), aYM = T(p.alpha), eps0 = T(p.eps0);
    const T dx = T(p.dx), dz = T(p.dz);
    const int tx = threadIdx.x % MX, tz = threadIdx.x / MX;

    // Ampere at a tile position (ix, iz) -> LDS; returns the real cell.
    auto ampere = [&](int ix, int iz) {
        int xr = (ix < ncol) ? x0 + ix : xp_src(x0 + ncol - 1, p);
        int zr = (iz < nrow) ? z0 + iz : zp_per(z0 + nrow - 1, nz);
        const size_t i = (size_t)xr + (size_t)zr * nx;
        const int zm = zm_per(zr, nz);
        const int xm = p.periodic_x ? (xr + nx - 1) % nx : (xr == 0 ? 0 : xr - 1);
        const size_t izm = (size_t)xr + (size_t)zm * nx, ixm = (size_t)xm + (size_t)zr * nx;
        T Jx[3], Jz[3];
        currents_at(O, xr, zr, p, Jx, Jz);
        const bool wall = !p.periodic_x && (xr == 0 || xr == nx - 1);   // YM_Fluid.cu:223-226
        T By[3], Byzm[3], Byxm[3], Az[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            By[c] = S.a[S_BY1 + c][i];
            Byzm[c] = S.a[S_BY1 + c][izm];
            Byxm[c] = S.a[S_BY1 + c][ixm];
            Az[c] = S.a[S_AZ1 + c][i];
        }
        T cr[3] = {Az[1] * By[2] - Az[2] * By[1],
                   Az[2] * By[0] - Az[0] * By[2],
                   Az[0] * By[1] - Az[1] * By[0]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T dBdz = (By[c] - Byzm[c]) / dz;
            T ex = S.a[S_EX1 + c][i];
            sEx[c][iz][ix] = ex + dt * (c2 * (-dBdz - aYM * cr[c]) - Jx[c] / eps0);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T dBdx = (By[c] - Byxm[c]) / dx;
            T ez = wall ? T(0.0) : S.a[S_EZ1 + c][i];
            sEz[c][iz][ix] = ez + dt * (c2 * dBdx - Jz[c] / eps0);
        }
    };

    // own positions: (tx, tz + MZT*k); halo: column ncol and row nrow
    for (int k = 0; k < MZ / MZT; ++k) {
        int iz = tz + MZT * k;
        if (tx < ncol && iz < nrow) ampere(tx, iz);
    }
    {
        int h = threadIdx.x;                      // 0 .. MX*MZT-1
        if (h < nrow + 1) ampere(ncol, h);        // halo column (incl. corner)
        else if (h - (nrow + 1) < ncol) ampere(h - (nrow + 1), nrow);   // halo row
    }
    __syncthreads();

    for (int k = 0; k < MZ / MZT; ++k) {
        int iz = tz + MZT * k;
        int ix = tx;
        if (ix >= ncol || iz >= nrow) continue;
        const int xr = x0 + ix, zr = z0 + iz;
        const size_t i = (size_t)xr + (size_t)zr * nx;
        // kernel_ym_potential (YM_Physics.cu:589-599)
        T Az[3], Ex[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Az[c] = S.a[S_AZ1 + c][i];
            if (c > 0 || !p.freeze_az1) Az[c] = Az[c] - dt * sEz[c][iz][ix];
            Ex[c] = sEx[c][iz][ix];
        }
        // kernel_ym_faraday (YM_Physics.cu:120-151).  Walled x: xp clamps to
        // nx-1, which the halo column holds when the tile ends at the wall.
        int ixp = ix + 1;
        if (!p.periodic_x && xr == nx - 1) ixp = ix;
        T cr[3] = {Az[1] * Ex[2] - Az[2] * Ex[1],
                   Az[2] * Ex[0] - Az[0] * Ex[2],
                   Az[0] * Ex[1] - Az[1] * Ex[0]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T dEx = (sEx[c][iz + 1][ix] - Ex[c]) / dz;
            T dEz = (sEz[c][iz][ixp] - sEz[c][iz][ix]) / dx;
            T by = S.a[S_BY1 + c][i];
            O.a[S_BY1 + c][i] = by - dt * (dEx - dEz - aYM * cr[c]);
            O.a[S_EX1 + c][i] = Ex[c];
            O.a[S_EZ1 + c][i] = sEz[c][iz][ix];
            if (c > 0 || !p.freeze_az1) O.a[S_AZ1 + c][i] = Az[c];
        }
    }
}

// ==========================================================================
// launchers
// ==========================================================================
template <typename T>
void launch_fct_x(const State<T>& S, const Stage<T>& out, const KParams& p, int sp0, int nsp,
                  hipStream_t st) {
    dim3 grid((p.nx + XB - 1) / XB, p.nz, nsp);
    hipLaunchKernelGGL(k_fct_x<T>, grid, dim3(XB), 0, st, S, out, p, sp0);
}

template <typename T>
void launch_fct_z(const State<T>& S, const Stage<T>& X, const State<T>& O, const KParams& p,
                  int sp0, int nsp, hipStream_t st) {
    constexpr int TZ = 16;
    int nchunk = (p.nz + TZ - 1) / TZ;
    dim3 grid((p.nx + ZB_X - 1) / ZB_X, (nchunk + ZB_C - 1) / ZB_C, nsp);
    hipLaunchKernelGGL((k_fct_z<T, TZ>), grid, dim3(ZB_X * ZB_C), 0, st, S, X, O, p, sp0);
}

template <typename T>
void launch_maxwell(const State<T>& S, const State<T>& O, const KParams& p, hipStream_t st) {
    dim3 grid((p.nx + MX - 1) / MX, (p.nz + MZ - 1) / MZ);
    hipLaunchKernelGGL(k_maxwell<T>, grid, dim3(MX * MZT), 0, st, S, O, p);
}

template void launch_fct_x<double>(const State<double>&, const Stage<double>&, const KParams&, int, int, hipStream_t);
template void launch_fct_x<float>(const State<float>&, const Stage<float>&, const KParams&, int, int, hipStream_t);
template void launch_fct_z<double>(const State<double>&, const Stage<double>&, const State<double>&, const KParams&, int, int, hipStream_t);
template void launch_fct_z<float>(const State<float>&, const Stage<float>&, const State<float>&, const KParams&, int, int, hipStream_t);
template void launch_maxwell<double>(const State<double>&, const State<double>&, const KParams&, hipStream_t);
template void launch_maxwell<float>(const State<float>&, const State<float>&, const KParams&, hipStream_t);

}  // namespace khs
