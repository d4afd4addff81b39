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
#include <algorithm>
#include <atomic>
#include <type_traits>

#include "khs_device.h"
#include "khs_launch.h"

namespace khs {

// ==========================================================================
// x sweep
// ==========================================================================
// XB cells per row segment (= threads), XW = segment + 3-cell halo each side.
// XB is tuned per grid (64 / 128 / 256): smaller segments use less LDS and
// so fit more workgroups per CU, which hides the load latency of phase A.
// VS = 1 runs phases C-E one variable at a time through single LOR / LNR /
// flux rows (less LDS, more workgroups per CU, 3 barriers per variable);
// VS = 0 keeps all variables' rows.  M is the model policy (YMModel /
// EMModel): variables, slots, sources.
template <typename T, int XB, typename M, int VS>
__global__ __launch_bounds__(XB) void k_fct_x(State<T> S, Stage<T> out, KParams p, int sp0, int nsp,
                                              int ntx, int rows) {
    constexpr int XW = XB + 6;
    constexpr int NV = M::NV;
    constexpr int NR = VS ? 1 : NV;   // LOR / LNR rows held at once
    __shared__ T sU[XW];
    __shared__ T sS[M::NXS][XW];
    __shared__ T sE[XW], sNU[XW], sMU[XW];
    __shared__ T sV[NV][XW];      // variable, later (VS = 0) the limited interface flux
    __shared__ T sLOR[NR][XW];
    __shared__ T sLNR[NR][XW];
    __shared__ T sFL[VS ? XW : 1];

    const int tid = threadIdx.x;
    // logical tile: species fastest, then x segments, then groups of `rows`
    // rows (one block sweeps its rows in turn: fewer, longer-lived
    // workgroups; row z-1, whose fields the x source reads, is in cache)
    const int nrg = (p.nz + rows - 1) / rows;
    int lt = xcd_logical(blockIdx.x, gridDim.x);
    if (lt >= nsp * ntx * nrg) return;
    lt = tile_order(lt, nsp * ntx * nrg, p.rev);
    const int sp = sp0 + lt % nsp;
    lt /= nsp;
    const int x0 = (lt % ntx) * XB;
    const int zb = (lt / ntx) * rows;
    const int ze = min(zb + rows, p.nz);
    const int b = M::base(sp);
    const T dt = T(p.dt);
    const T dtdx = T(p.dtdx);
    for (int z = zb; z < ze; ++z) {
    const size_t row = (size_t)z * p.nx;

    // Phase A: stage variables, velocity and x sources for positions
    // x0-3 .. x0+XB+2 (position q <-> extended x = x0-3+q).
    for (int q = tid; q < XW; q += XB) {
        const int xr = map_x(x0 - 3 + q, p);
        const unsigned i = (unsigned)row + (unsigned)xr;
        T v[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] = ld32(S.a[b + k], i);
        T n_safe = kmax(v[0], T(0.05));
        sU[q] = v[M::IPX] / n_safe;                         // update_velocities
        T src[M::NXS];
        M::x_src(S, sp, xr, z, p, src);
#pragma unroll
        for (int k = 0; k < M::NXS; ++k) sS[k][q] = src[k];
#pragma unroll
        for (int k = 0; k < NV; ++k) sV[k][q] = v[k];
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

    if constexpr (VS == 0) {
    // Phase C: transported (LORHOT) and diffused (LNRHOT) low-order values at
    // positions 1 .. XW-2 (FCT_Sweeps.cu:88-103).
    for (int q = 1 + tid; q < XW - 1; q += XB) {
        const T eL = sE[q - 1], eR = sE[q];
        const T nL = sNU[q - 1], nR = sNU[q];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const T vm = sV[v][q - 1], v0 = sV[v][q], vp = sV[v][q + 1];
            T fL = T(0.5) * eL * (vm + v0);
            T fR = T(0.5) * eR * (v0 + vp);
            T dL = nL * (v0 - vm);
            T dR = nR * (vp - v0);
            const int sl = M::xslot(v);
            T src = (sl >= 0) ? sS[sl < 0 ? 0 : sl][q] : T(0.0);
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
        for (int v = 0; v < NV; ++v) {
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
        for (int v = 0; v < NV; ++v) {
            T r = sLNR[v][q] - (sV[v][q] - sV[v][q - 1]);
            st32_nt(out.a[sp * NV + v], (unsigned)row + (unsigned)x, r);
        }
    }
    } else {
    // the same phases C, D, E one variable at a time
    const int x = x0 + tid;
#pragma unroll 1
    for (int v = 0; v < NV; ++v) {
        const int sl = M::xslot(v);
        for (int q = 1 + tid; q < XW - 1; q += XB) {
            const T eL = sE[q - 1], eR = sE[q];
            const T nL = sNU[q - 1], nR = sNU[q];
            const T vm = sV[v][q - 1], v0 = sV[v][q], vp = sV[v][q + 1];
            T fL = T(0.5) * eL * (vm + v0);
            T fR = T(0.5) * eR * (v0 + vp);
            T dL = nL * (v0 - vm);
            T dR = nR * (vp - v0);
            T src = (sl >= 0) ? sS[sl < 0 ? 0 : sl][q] : T(0.0);
            T lor = v0 + dt * src - (fR - fL);
            sLOR[0][q] = lor;
            sLNR[0][q] = lor + (dR - dL);
        }
        __syncthreads();
        for (int q = 2 + tid; q < XW - 3; q += XB) {
            T flx = sMU[q] * (sLOR[0][q + 1] - sLOR[0][q]);
            T l0 = sLNR[0][q - 1], l1 = sLNR[0][q], l2 = sLNR[0][q + 1], l3 = sLNR[0][q + 2];
            sFL[q] = fct_limit(flx, l1 - l0, l2 - l1, l3 - l2);
        }
        __syncthreads();
        if (x < p.nx) {
            const int q = 3 + tid;
            st32_nt(out.a[sp * NV + v], (unsigned)row + (unsigned)x, sLNR[0][q] - (sFL[q] - sFL[q - 1]));
        }
        __syncthreads();              // sLOR / sLNR / sFL are rewritten for the next variable
    }
    }
    __syncthreads();                  // LDS is restaged for the next row
    }
}

// ==========================================================================
// x sweep, register form
// ==========================================================================
// One wavefront sweeps a row segment with no LDS and no barriers.  Lane l owns
// the C consecutive cells xl = x0 - C + C*l .. xl+C-1 in registers and takes
// the values the FCT stencil needs across its run boundaries from lanes l-1 /
// l+1 by whole-wave DPP shifts.  Lanes 0 and 63 are halo lanes: the stencil
// reaches 3 cells, so with C >= 4 lane 1 needs only what lane 0 computes from
// its own cells 1..3 (and lane 62 only lane 63's cells 0..2); their outer
// values, fed by the shifts' zero fill, are never used.  The wave writes the
// 62*C cells of lanes 1..62.  Every quantity is evaluated at the same extended
// positions (map_x) and with the same expressions as in k_fct_x, so the two
// forms agree bit for bit.  Lanes whose run is inside the row load and store
// with 16-byte vectors when rows are C-aligned; a lane with cells outside the
// row (domain ends) loads each cell through map_x and evaluates its px source
// cell by cell (the x-1 / x+1 neighbours of such cells follow xm_src / xp_src
// of the mapped cell, not the lane shift).  YM model only.
constexpr int XR_OUT_LANES = 62;

template <typename T, int C>
__device__ __forceinline__ void load_run(const T* a, unsigned row, int xl, bool vec, const KParams& p,
                                         T (&o)[C]) {
    if (vec) {
        constexpr int N = 16 / sizeof(T);
        typedef T V __attribute__((ext_vector_type(N)));
        const V* q = reinterpret_cast<const V*>(reinterpret_cast<const char*>(a) +
                                                (row + (unsigned)xl) * (unsigned)sizeof(T));
#pragma unroll
        for (int k = 0; k < C / N; ++k) {
            const V w = q[k];
#pragma unroll
            for (int j = 0; j < N; ++j) o[k * N + j] = w[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < C; ++j) o[j] = ld32(a, row + (unsigned)map_x(xl + j, p));
    }
}

template <typename T, int C>
__device__ __forceinline__ void store_run(T* a, unsigned row, int xl, bool vec, int nx, const T (&o)[C]) {
    if (vec) {
        constexpr int N = 16 / sizeof(T);
        typedef T V __attribute__((ext_vector_type(N)));
        V* q = reinterpret_cast<V*>(reinterpret_cast<char*>(a) + (row + (unsigned)xl) * (unsigned)sizeof(T));
#pragma unroll
        for (int k = 0; k < C / N; ++k) {
            V w;
#pragma unroll
            for (int j = 0; j < N; ++j) w[j] = o[k * N + j];
            __builtin_nontemporal_store(w, q + k);
        }
    } else {
#pragma unroll
        for (int j = 0; j < C; ++j)
            if (xl + j < nx) st32_nt(a, row + (unsigned)(xl + j), o[j]);
    }
}

template <typename T, int C>
__global__ __launch_bounds__(256) void k_fct_xr(State<T> S, Stage<T> out, KParams p, int sp0, int nsp, int nseg) {
    static_assert(C >= 4 && C % (16 / sizeof(T)) == 0, "run length");
    constexpr int NV = F_COUNT;
    const int lane = threadIdx.x & 63;
    int wl = xcd_logical(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (wl >= nsp * nseg * p.nz) return;       // whole wavefronts leave together
    wl = tile_order(wl, nsp * nseg * p.nz, p.rev);
    const int sp = sp0 + wl % nsp;
    const int seg = (wl / nsp) % nseg;
    const int z = wl / (nsp * nseg);
    const int nx = p.nx;
    const int xl = seg * (XR_OUT_LANES * C) - C + C * lane;
    const unsigned row = (unsigned)z * (unsigned)nx;
    const unsigned rowm = (unsigned)zm_per(z, p.nz) * (unsigned)nx;
    const bool inside = xl >= 0 && xl + C <= nx;
    const bool vec = inside && (nx % C) == 0;
    const int b = YMModel::base(sp);
    const T dt = T(p.dt);
    const T dtdx = T(p.dtdx);

    T v[NV][C];
#pragma unroll
    for (int k = 0; k < NV; ++k) load_run<T, C>(S.a[b + k], row, xl, vec, p, v[k]);

    // update_velocities and the px Lorentz source (kernel_ym_lorentz,
    // YM_Fluid.cu:32-84) of every cell of the run
    T u[C], src[C];
#pragma unroll
    for (int j = 0; j < C; ++j) u[j] = v[F_PX][j] / kmax(v[F_N][j], T(0.05));
    {
        T acc[3][C];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T ex[C], by[C], bzm[C];
            load_run<T, C>(S.a[S_EX1 + c], row, xl, vec, p, ex);
            load_run<T, C>(S.a[S_BY1 + c], row, xl, vec, p, by);
            load_run<T, C>(S.a[S_BY1 + c], rowm, xl, vec, p, bzm);
            const T exl = lane_from_left(ex[C - 1]);
            const T byl = lane_from_left(by[C - 1]);
            const T bzml = lane_from_left(bzm[C - 1]);
#pragma unroll
            for (int j = 0; j < C; ++j) {
                const T n_safe = kmax(v[F_N][j], T(0.05));
                const T vz = v[F_PZ][j] / n_safe;
                const T exm = j ? ex[j - 1] : exl;
                const T bym = j ? by[j - 1] : byl;
                const T bzmm = j ? bzm[j - 1] : bzml;
                T Exn = T(0.5) * (ex[j] + exm);
                T Byn = T(0.25) * (by[j] + bym + bzm[j] + bzmm);
                acc[c][j] = v[F_Q1 + c][j] * (Exn + vz * Byn);
            }
        }
#pragma unroll
        for (int j = 0; j < C; ++j) src[j] = -(acc[0][j] + acc[1][j] + acc[2][j]);
        if (p.warm_T > 0.0) {
            const T nl = lane_from_left(v[F_N][C - 1]);
            const T nr = lane_from_right(v[F_N][0]);
#pragma unroll
            for (int j = 0; j < C; ++j) {
                const T nm = j ? v[F_N][j - 1] : nl;
                const T np = (j + 1 < C) ? v[F_N][j + 1] : nr;
                T dndx = (np - nm) / (T(2.0) * T(p.dx));
                src[j] -= T(p.warm_T) * dndx;
            }
        }
    }
    if (!inside) {
#pragma unroll
        for (int j = 0; j < C; ++j) src[j] = lorentz_px(S, sp, map_x(xl + j, p), z, p);
    }

    // interface j+1/2 coefficients (compute_eps, compute_nu_mu)
    T eps[C], nu[C], mu[C];
    {
        const T ur = lane_from_right(u[0]);
#pragma unroll
        for (int j = 0; j < C; ++j) {
            T vi = T(0.5) * (u[j] + ((j + 1 < C) ? u[j + 1] : ur));
            eps[j] = vi * dtdx;
            fct_nu_mu(eps[j], nu[j], mu[j]);
        }
    }
    const T epsl = lane_from_left(eps[C - 1]);
    const T nul = lane_from_left(nu[C - 1]);
    const bool store = lane >= 1 && lane <= XR_OUT_LANES;

#pragma unroll
    for (int k = 0; k < NV; ++k) {
        // LORHOT / LNRHOT (FCT_Sweeps.cu:88-103)
        const T vl = lane_from_left(v[k][C - 1]);
        const T vr = lane_from_right(v[k][0]);
        T lor[C], lnr[C];
#pragma unroll
        for (int j = 0; j < C; ++j) {
            const T vm = j ? v[k][j - 1] : vl, v0 = v[k][j], vp = (j + 1 < C) ? v[k][j + 1] : vr;
            const T eL = j ? eps[j - 1] : epsl, eR = eps[j];
            const T nL = j ? nu[j - 1] : nul, nR = nu[j];
            T fL = T(0.5) * eL * (vm + v0);
            T fR = T(0.5) * eR * (v0 + vp);
            T dL = nL * (v0 - vm);
            T dR = nR * (vp - v0);
            T sj = (k == F_PX) ? src[j] : T(0.0);
            T lo = v0 + dt * sj - (fR - fL);
            lor[j] = lo;
            lnr[j] = lo + (dR - dL);
        }
        // limited antidiffusive flux at interfaces j+1/2 (FCT_Sweeps.cu:139, 155-160)
        const T lorr = lane_from_right(lor[0]);
        const T lnrl = lane_from_left(lnr[C - 1]);
        const T lnrr = lane_from_right(lnr[0]);
        const T lnrr2 = lane_from_right(lnr[1]);
        T fl[C];
#pragma unroll
        for (int j = 0; j < C; ++j) {
            T flx = mu[j] * (((j + 1 < C) ? lor[j + 1] : lorr) - lor[j]);
            const T l0 = j ? lnr[j - 1] : lnrl;
            const T l1 = lnr[j];
            const T l2 = (j + 1 < C) ? lnr[j + 1] : lnrr;
            const T l3 = (j + 2 < C) ? lnr[j + 2] : ((j + 2 == C) ? lnrr : lnrr2);
            fl[j] = fct_limit(flx, l1 - l0, l2 - l1, l3 - l2);
        }
        const T fll = lane_from_left(fl[C - 1]);
        T r[C];
#pragma unroll
        for (int j = 0; j < C; ++j) r[j] = lnr[j] - (fl[j] - (j ? fl[j - 1] : fll));
        if (store) store_run<T, C>(out.a[sp * NV + k], row, xl, vec, nx, r);
    }
}

// ==========================================================================
// z sweep (register march)
// ==========================================================================
// FCT pipeline along z for one column, kept in registers.  push() is fed the
// extended rows e = z0-3 .. zend+2 in order (rows outside [0, nz) are the
// clamped end rows, FCT_Sweeps.cu:52,73,78) and yields row e-3 once
// e >= z0+3.  Per row it evaluates the interface (e-1, e) coefficients, the
// transport/diffusion fluxes, LORHOT/LNRHOT of row e-1, the antidiffusive flux
// at e-3/2 and the limited flux at e-5/2, in the reference's order
// (FCT_Sweeps.cu:88-162).
template <typename T, int NV = F_COUNT>
struct ZPipe {
    T Vm[NV], Sm[NV];                // variable and source at row e-1
    T Fm[NV], Dm[NV];                // transport / diffusion flux at e-3/2
    T LORm[NV];                      // LORHOT(e-2)
    T L3[NV], L2[NV], L1[NV];        // LNRHOT(e-4), (e-3), (e-2)
    T FLXm[NV];                      // antidiffusive flux at e-5/2
    T FLm[NV];                       // limited flux at e-7/2
    T Um, MUm;                       // velocity at e-1, mu at interface e-3/2

    __device__ __forceinline__ bool push(int e, int z0, const T (&V)[NV], T u,
                                         const T (&Sv)[NV], T dt, T dtdz, T (&out)[NV]) {
        bool have = false;
        if (e > z0 - 3) {
            T vi = T(0.5) * (Um + u);
            T eps = vi * dtdz;
            T nu, mu;
            fct_nu_mu(eps, nu, mu);
            const bool do_lor = (e >= z0 - 1);
            const bool do_flx = (e >= z0);
            const bool do_fl = (e >= z0 + 2);
            have = (e >= z0 + 3);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                T F = T(0.5) * eps * (Vm[v] + V[v]);
                T D = nu * (V[v] - Vm[v]);
                if (do_lor) {
                    T lor = Vm[v] + dt * Sm[v] - (F - Fm[v]);        // row e-1
                    T lnr = lor + (D - Dm[v]);
                    T flx = T(0);
                    if (do_flx) flx = MUm * (lor - LORm[v]);         // interface e-3/2
                    if (do_fl) {
                        T fl = fct_limit(FLXm[v], L2[v] - L3[v], L1[v] - L2[v], lnr - L1[v]);  // e-5/2
                        out[v] = L2[v] - (fl - FLm[v]);               // row e-3
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
        for (int v = 0; v < NV; ++v) { Vm[v] = V[v]; Sm[v] = Sv[v]; }
        Um = u;
        return have;
    }
    // push() for a row e >= z0 + 3, where every stage is active: the same
    // expressions without the warm-up conditions, so that an unrolled march
    // keeps the pipeline in renamed registers instead of moving it.
    __device__ __forceinline__ void push_steady(const T (&V)[NV], T u, const T (&Sv)[NV], T dt, T dtdz,
                                                T (&out)[NV]) {
        T vi = T(0.5) * (Um + u);
        T eps = vi * dtdz;
        T nu, mu;
        fct_nu_mu(eps, nu, mu);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            T F = T(0.5) * eps * (Vm[v] + V[v]);
            T D = nu * (V[v] - Vm[v]);
            T lor = Vm[v] + dt * Sm[v] - (F - Fm[v]);
            T lnr = lor + (D - Dm[v]);
            T flx = MUm * (lor - LORm[v]);
            T fl = fct_limit(FLXm[v], L2[v] - L3[v], L1[v] - L2[v], lnr - L1[v]);
            out[v] = L2[v] - (fl - FLm[v]);
            FLm[v] = fl;
            FLXm[v] = flx;
            LORm[v] = lor;
            L3[v] = L2[v]; L2[v] = L1[v]; L1[v] = lnr;
            Fm[v] = F; Dm[v] = D;
        }
        MUm = mu;
#pragma unroll
        for (int v = 0; v < NV; ++v) { Vm[v] = V[v]; Sm[v] = Sv[v]; }
        Um = u;
    }
};

// Inputs of the z sweep at real cell (x, rc) from the step-start state: the
// stale z velocity (YM_FCT.cu:143-146 transposes d_old.uy, set before the
// x sweep) and the pz / Q^a sources of the previous step's end.
template <typename T>
__device__ __forceinline__ void z_inputs(const State<T>& S, int sp, int x, int rc, const KParams& p,
                                         T& u, T (&Sv)[F_COUNT]) {
    const int b = S_NA + sp * F_COUNT;
    const size_t i = (size_t)x + (size_t)rc * p.nx;
    T n_safe = kmax(S.a[b + F_N][i], T(0.05));
    u = S.a[b + F_PZ][i] / n_safe;
    Sv[F_N] = T(0);
    Sv[F_PX] = T(0);
    Sv[F_PZ] = lorentz_pz(S, sp, x, rc, p);
    T q[3];
    precession_at(S, sp, x, rc, p, q);
    Sv[F_Q1] = q[0]; Sv[F_Q2] = q[1]; Sv[F_Q3] = q[2];
}

// Swept row written to the next state with the post-sweep boundary condition
// of kernel_enforce_boundaries: density floor and x-wall px (YM_Fluid.cu:13-30).
template <typename T, int NV, int IPX>
__device__ __forceinline__ void store_fluid(const State<T>& O, int b, size_t io, bool wall, T (&r)[NV]) {
    if (r[0] < T(0.05)) r[0] = T(0.05);
    if (wall) r[IPX] = T(0.0);
#pragma unroll
    for (int v = 0; v < NV; ++v) O.a[b + v][io] = r[v];
}

constexpr int ZB_X = 64;           // columns per workgroup (one wave wide)
constexpr int ZB_C = 4;            // z chunks per workgroup (one per wave)

// One lane marches one column of one species down a chunk of TZ rows (plus
// the 3-row FCT halo on each side).  The field values that the z sources
// read at row z-1 are carried over from the previous row in registers, so
// every input array is loaded once per row.  PF = 1 issues the loads of row
// e+1 before the arithmetic of row e (software pipelining); PF = 0 loads each
// row when it is used (fewer registers).
template <typename T, int TZ, int PF, typename M>
__global__ __launch_bounds__(ZB_X * ZB_C) void k_fct_z(State<T> S, Stage<T> X, State<T> O,
                                                       KParams p, int sp0, int nsp, int ntx, int ncg) {
    constexpr int NV = M::NV;
    typedef typename M::template Row<T> Row;
    // logical tile: species fastest (both species read the same fields for
    // their sources), then 64-column groups, then groups of ZB_C z chunks
    int lt = xcd_logical(blockIdx.x, gridDim.x);
    if (lt >= nsp * ntx * ncg) return;
    lt = tile_order(lt, nsp * ntx * ncg, p.rev);
    const int sp = sp0 + lt % nsp;
    lt /= nsp;
    const int x = (lt % ntx) * ZB_X + (threadIdx.x % ZB_X);
    const int chunk = (lt / ntx) * ZB_C + (threadIdx.x / ZB_X);
    const int z0 = chunk * TZ;
    if (x >= p.nx || z0 >= p.nz) return;
    const int nx = p.nx, nz = p.nz;
    const int b = M::base(sp);
    const int xm = xm_src(x, p);
    const T dt = T(p.dt), dtdz = T(p.dtdz);
    const int zend = min(z0 + TZ, nz);
    const int elast = zend + 2;
    const bool wall = !p.periodic_x && (x == 0 || x == nx - 1);

    auto load = [&](int r, Row& row, T (&F)[NV], T (&V)[NV]) {
        const unsigned i = (unsigned)x + (unsigned)r * (unsigned)nx;
        M::load_row(S, i, (unsigned)xm + (unsigned)r * (unsigned)nx, row);
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v] = ld32(S.a[b + v], i);
#pragma unroll
        for (int v = 0; v < NV; ++v) V[v] = ld32(X.a[sp * NV + v], i);
    };

    // rows outside [0, nz) are the clamped end rows (FCT_Sweeps.cu:52,73,78)
    int rc = clampi(z0 - 3, 0, nz - 1);
    Row prev, cur;
    T F[NV], V[NV];
    {
        T Fd[NV], Vd[NV];
        load(zm_per(rc, nz), prev, Fd, Vd);
    }
    load(rc, cur, F, V);
    ZPipe<T, NV> zp;
    zp.Um = T(0);
    zp.MUm = T(0);
    for (int e = z0 - 3; e <= elast; ++e) {
        const int rn = clampi(min(e + 1, elast), 0, nz - 1);
        Row nrow;
        T nF[NV], nV[NV];
        if (PF) load(rn, nrow, nF, nV);
        T u, Sv[NV], out[NV];
        M::z_src_rows(F, cur, prev, S, sp, x, rc, p, u, Sv);
        if (zp.push(e, z0, V, u, Sv, dt, dtdz, out)) {
            if (out[0] < T(0.05)) out[0] = T(0.05);          // kernel_enforce_boundaries
            if (wall) out[M::IPX] = T(0.0);
            const unsigned io = (unsigned)x + (unsigned)(e - 3) * (unsigned)nx;
#pragma unroll
            for (int v = 0; v < NV; ++v) st32_nt(O.a[b + v], io, out[v]);
        }
        if (!PF) load(rn, nrow, nF, nV);
        // advance: row z-1 of the next row is this row unless the clamp holds
        // the row index at an end
        M::sel_row(rn != rc, cur, prev);
        cur = nrow;
#pragma unroll
        for (int v = 0; v < NV; ++v) { F[v] = nF[v]; V[v] = nV[v]; }
        rc = rn;
    }
}

// The same z march with the loads issued two rows ahead: at occupancy one
// (the pipeline state of six variables fills the register file) a one-row
// prefetch leaves part of the load latency exposed; the extra row of inputs
// lives in the otherwise unused accumulation registers.
template <typename T, int TZ, typename M>
__global__ __launch_bounds__(ZB_X * ZB_C) void k_fct_z2(State<T> S, Stage<T> X, State<T> O,
                                                        KParams p, int sp0, int nsp, int ntx, int ncg) {
    constexpr int NV = M::NV;
    typedef typename M::template Row<T> Row;
    int lt = xcd_logical(blockIdx.x, gridDim.x);
    if (lt >= nsp * ntx * ncg) return;
    lt = tile_order(lt, nsp * ntx * ncg, p.rev);
    const int sp = sp0 + lt % nsp;
    lt /= nsp;
    const int x = (lt % ntx) * ZB_X + (threadIdx.x % ZB_X);
    const int chunk = (lt / ntx) * ZB_C + (threadIdx.x / ZB_X);
    const int z0 = chunk * TZ;
    if (x >= p.nx || z0 >= p.nz) return;
    const int nx = p.nx, nz = p.nz;
    const int b = M::base(sp);
    const int xm = xm_src(x, p);
    const T dt = T(p.dt), dtdz = T(p.dtdz);
    const int zend = min(z0 + TZ, nz);
    const int elast = zend + 2;
    const bool wall = !p.periodic_x && (x == 0 || x == nx - 1);

    auto load = [&](int r, Row& row, T (&F)[NV], T (&V)[NV]) {
        const unsigned i = (unsigned)x + (unsigned)r * (unsigned)nx;
        M::load_row(S, i, (unsigned)xm + (unsigned)r * (unsigned)nx, row);
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v] = ld32(S.a[b + v], i);
#pragma unroll
        for (int v = 0; v < NV; ++v) V[v] = ld32(X.a[sp * NV + v], i);
    };
    auto row_at = [&](int e) { return clampi(min(e, elast), 0, nz - 1); };

    int rc = row_at(z0 - 3);
    Row prev, cur, nrow;
    T F[NV], V[NV], nF[NV], nV[NV];
    {
        T Fd[NV], Vd[NV];
        load(zm_per(rc, nz), prev, Fd, Vd);
    }
    load(rc, cur, F, V);
    int rn = row_at(z0 - 2);
    load(rn, nrow, nF, nV);
    ZPipe<T, NV> zp;
    zp.Um = T(0);
    zp.MUm = T(0);
    for (int e = z0 - 3; e <= elast; ++e) {
        const int rnn = row_at(e + 2);
        Row nnrow;
        T nnF[NV], nnV[NV];
        load(rnn, nnrow, nnF, nnV);
        T u, Sv[NV], out[NV];
        M::z_src_rows(F, cur, prev, S, sp, x, rc, p, u, Sv);
        if (zp.push(e, z0, V, u, Sv, dt, dtdz, out)) {
            if (out[0] < T(0.05)) out[0] = T(0.05);          // kernel_enforce_boundaries
            if (wall) out[M::IPX] = T(0.0);
            const unsigned io = (unsigned)x + (unsigned)(e - 3) * (unsigned)nx;
#pragma unroll
            for (int v = 0; v < NV; ++v) st32_nt(O.a[b + v], io, out[v]);
        }
        M::sel_row(rn != rc, cur, prev);
        cur = nrow;
        nrow = nnrow;
#pragma unroll
        for (int v = 0; v < NV; ++v) { F[v] = nF[v]; V[v] = nV[v]; nF[v] = nnF[v]; nV[v] = nnV[v]; }
        rc = rn;
        rn = rnn;
    }
}

// The same z march with the variables split into two groups (model's
// NG = NV/2 each) on two waves: half the pipeline state per lane, so more
// waves fit; both waves of a column read the cell's fluid for the sources
// (the second read hits L2).  A workgroup is 64 columns x 2 chunks x 2 groups.
template <typename T, int TZ, int PF, typename M, int G>
__device__ __forceinline__ void z_march_g(const State<T>& S, const Stage<T>& X, const State<T>& O,
                                          const KParams& p, int sp, int x, int z0) {
    constexpr int NV = M::NV, NG = M::NG;
    typedef typename M::template Row<T> Row;
    const int nx = p.nx, nz = p.nz;
    const int b = M::base(sp);
    const int xm = xm_src(x, p);
    const T dt = T(p.dt), dtdz = T(p.dtdz);
    const int zend = min(z0 + TZ, nz);
    const int elast = zend + 2;
    const bool wall = !p.periodic_x && (x == 0 || x == nx - 1);

    auto load = [&](int r, Row& row, T (&F)[NV], T (&V)[NG]) {
        const unsigned i = (unsigned)x + (unsigned)r * (unsigned)nx;
        M::template load_row_g<G>(S, i, (unsigned)xm + (unsigned)r * (unsigned)nx, row);
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v] = ld32(S.a[b + v], i);
#pragma unroll
        for (int j = 0; j < NG; ++j) V[j] = ld32(X.a[sp * NV + G * NG + j], i);
    };

    int rc = clampi(z0 - 3, 0, nz - 1);
    Row prev, cur;
    T F[NV], V[NG];
    {
        T Fd[NV], Vd[NG];
        load(zm_per(rc, nz), prev, Fd, Vd);
    }
    load(rc, cur, F, V);
    ZPipe<T, NG> zp;
    zp.Um = T(0);
    zp.MUm = T(0);
    for (int e = z0 - 3; e <= elast; ++e) {
        const int rn = clampi(min(e + 1, elast), 0, nz - 1);
        Row nrow;
        T nF[NV], nV[NG];
        if (PF) load(rn, nrow, nF, nV);
        T u, Sv[NG], out[NG];
        M::template z_src_rows_g<G>(F, cur, prev, S, sp, x, rc, p, u, Sv);
        if (zp.push(e, z0, V, u, Sv, dt, dtdz, out)) {
            if (G == 0) {
                if (out[0] < T(0.05)) out[0] = T(0.05);      // kernel_enforce_boundaries
                if (wall) out[M::IPX] = T(0.0);
            }
            const unsigned io = (unsigned)x + (unsigned)(e - 3) * (unsigned)nx;
#pragma unroll
            for (int j = 0; j < NG; ++j) st32(O.a[b + G * NG + j], io, out[j]);
        }
        if (!PF) load(rn, nrow, nF, nV);
        M::template sel_row_g<G>(rn != rc, cur, prev);
        cur = nrow;
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v] = nF[v];
#pragma unroll
        for (int j = 0; j < NG; ++j) V[j] = nV[j];
        rc = rn;
    }
}

template <typename T, int TZ, int PF, typename M>
__global__ __launch_bounds__(256) void k_fct_zg(State<T> S, Stage<T> X, State<T> O, KParams p, int sp0, int nsp,
                                                int ntx, int ncg) {
    int lt = xcd_logical(blockIdx.x, gridDim.x);
    if (lt >= nsp * ntx * ncg) return;
    lt = tile_order(lt, nsp * ntx * ncg, p.rev);
    const int sp = sp0 + lt % nsp;
    lt /= nsp;
    const int x = (lt % ntx) * 64 + (int)(threadIdx.x & 63);
    const int w = (int)(threadIdx.x >> 6);
    const int z0 = ((lt / ntx) * 2 + (w >> 1)) * TZ;
    if (x >= p.nx || z0 >= p.nz) return;
    if ((w & 1) == 0) z_march_g<T, TZ, PF, M, 0>(S, X, O, p, sp, x, z0);
    else z_march_g<T, TZ, PF, M, 1>(S, X, O, p, sp, x, z0);
}

// ==========================================================================
// fused x + z sweep
// ==========================================================================
// One wave owns 64 columns of one species and marches down a z chunk.  For
// every extended row it stages the row segment (+3 halo) in LDS, performs the
// x sweep there (the k_fct_x phases) and feeds its own column's x-swept values
// straight into the register z pipeline, so the intermediate fluid never goes
// to HBM.  The x sweep is recomputed for the 6 halo rows of each chunk.
// Blocks are numbered XCD-aware: consecutive logical tiles (species pair, then
// x tiles, then z chunks) land on the same XCD and share its L2.
constexpr int XZ_W = 64;
constexpr int XZ_P = XZ_W + 6;

template <typename T, int TZ>
__global__ __launch_bounds__(XZ_W) void k_fct_xz(State<T> S, State<T> O, KParams p, int ntx, int nch) {
    __shared__ T sU[XZ_P], sS[XZ_P], sE[XZ_P], sNU[XZ_P], sMU[XZ_P];
    __shared__ T sV[F_COUNT][XZ_P];
    __shared__ T sLOR[F_COUNT][XZ_P];
    __shared__ T sLNR[F_COUNT][XZ_P];

    int lb = xcd_logical(blockIdx.x, gridDim.x);
    if (lb >= 2 * ntx * nch) return;
    const int sp = lb & 1;
    lb >>= 1;
    const int x0 = (lb % ntx) * XZ_W;
    const int z0 = (lb / ntx) * TZ;
    const int lane = threadIdx.x;
    const int x = x0 + lane;
    const int nx = p.nx, nz = p.nz;
    if (z0 >= nz) return;                    // uniform per block
    const bool active = x < nx;
    const int b = S_NA + sp * F_COUNT;
    const T dt = T(p.dt), dtdx = T(p.dtdx), dtdz = T(p.dtdz);
    const int zend = min(z0 + TZ, nz);
    const bool wall = !p.periodic_x && (x == 0 || x == nx - 1);
    ZPipe<T> zp;
    zp.Um = T(0);
    zp.MUm = T(0);

    for (int e = z0 - 3; e <= zend + 2; ++e) {
        const int rc = clampi(e, 0, nz - 1);
        const size_t row = (size_t)rc * nx;
        // ---- x sweep of row rc, positions q <-> x = x0-3+q ----
        for (int q = lane; q < XZ_P; q += XZ_W) {
            const int xr = map_x(x0 - 3 + q, p);
            const size_t i = row + xr;
            T n = S.a[b + F_N][i];
            T px = S.a[b + F_PX][i];
            T n_safe = kmax(n, T(0.05));
            sU[q] = px / n_safe;
            sS[q] = lorentz_px(S, sp, xr, rc, p);
            sV[F_N][q] = n;
            sV[F_PX][q] = px;
#pragma unroll
            for (int v = F_PZ; v < F_COUNT; ++v) sV[v][q] = S.a[b + v][i];
        }
        __syncthreads();
        for (int q = lane; q < XZ_P - 1; q += XZ_W) {
            T vi = T(0.5) * (sU[q] + sU[q + 1]);
            T eps = vi * dtdx;
            T nu, mu;
            fct_nu_mu(eps, nu, mu);
            sE[q] = eps; sNU[q] = nu; sMU[q] = mu;
        }
        __syncthreads();
        for (int q = 1 + lane; q < XZ_P - 1; q += XZ_W) {
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
        for (int q = 2 + lane; q < XZ_P - 3; q += XZ_W) {
            const T mu = sMU[q];
#pragma unroll
            for (int v = 0; v < F_COUNT; ++v) {
                T flx = mu * (sLOR[v][q + 1] - sLOR[v][q]);
                T l0 = sLNR[v][q - 1], l1 = sLNR[v][q], l2 = sLNR[v][q + 1], l3 = sLNR[v][q + 2];
                sV[v][q] = fct_limit(flx, l1 - l0, l2 - l1, l3 - l2);
            }
        }
        __syncthreads();
        T V[F_COUNT];
        {
            const int q = 3 + lane;
#pragma unroll
            for (int v = 0; v < F_COUNT; ++v) V[v] = sLNR[v][q] - (sV[v][q] - sV[v][q - 1]);
        }
        __syncthreads();                     // LDS is restaged for the next row
        if (!active) continue;
        // ---- z pipeline ----
        T Sv[F_COUNT], u, out[F_COUNT];
        z_inputs(S, sp, x, rc, p, u, Sv);
        if (zp.push(e, z0, V, u, Sv, dt, dtdz, out))
            store_fluid<T, F_COUNT, F_PX>(O, b, (size_t)x + (size_t)(e - 3) * nx, wall, out);
    }
}

// ==========================================================================
// Maxwell: currents + Ampere + potential + Faraday
// ==========================================================================
constexpr int MX = 64, MZ = 16, MZT = 4;   // tile 64 x 16, 64 x 4 threads

template <typename T, int TM>
__global__ __launch_bounds__(MX * MZT) void k_maxwell(State<T> S, State<T> O, KParams p, int ntx, int ntz,
                                                      const double* decay, const int* action) {
    __shared__ T sEx[3][TM + 1][MX + 1];
    __shared__ T sEz[3][TM + 1][MX + 1];
    const int nx = p.nx, nz = p.nz;
    // logical tile: x tiles fastest (they share the x +- 1 halo lines)
    const int lt = tile_order(xcd_logical(blockIdx.x, gridDim.x), ntx * ntz, p.rev);
    if (lt >= ntx * ntz || lt < 0) return;
    const int x0 = (lt % ntx) * MX, z0 = (lt / ntx) * TM;
    const int ncol = min(MX, nx - x0), nrow = min(TM, nz - z0);
    const T dt = T(p.dt), c2 = T(p.c2), aYM = T(p.alpha), eps0 = T(p.eps0);
    const T dx = T(p.dx), dz = T(p.dz);
    const int tx = threadIdx.x % MX, tz = threadIdx.x / MX;

    // Ampere at a tile position (ix, iz) -> LDS; returns the real cell.
    auto ampere = [&](int ix, int iz) {
        int xr = (ix < ncol) ? x0 + ix : xp_src(x0 + ncol - 1, p);
        int zr = (iz < nrow) ? z0 + iz : zp_per(z0 + nrow - 1, nz);
        const unsigned i = (unsigned)xr + (unsigned)zr * (unsigned)nx;
        const int zm = zm_per(zr, nz);
        const int xm = p.periodic_x ? (xr + nx - 1) % nx : (xr == 0 ? 0 : xr - 1);
        const unsigned izm = (unsigned)xr + (unsigned)zm * (unsigned)nx, ixm = (unsigned)xm + (unsigned)zr * (unsigned)nx;
        T Jx[3], Jz[3];
        currents_at(O, xr, zr, p, Jx, Jz);
        const bool wall = !p.periodic_x && (xr == 0 || xr == nx - 1);   // YM_Fluid.cu:223-226
        T By[3], Byzm[3], Byxm[3], Az[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            By[c] = ld32(S.a[S_BY1 + c], i);
            Byzm[c] = ld32(S.a[S_BY1 + c], izm);
            Byxm[c] = ld32(S.a[S_BY1 + c], ixm);
            Az[c] = ld32(S.a[S_AZ1 + c], i);
        }
        T cr[3] = {Az[1] * By[2] - Az[2] * By[1],
                   Az[2] * By[0] - Az[0] * By[2],
                   Az[0] * By[1] - Az[1] * By[0]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T dBdz = (By[c] - Byzm[c]) / dz;
            T ex = ld32(S.a[S_EX1 + c], i);
            sEx[c][iz][ix] = ex + dt * (c2 * (-dBdz - aYM * cr[c]) - Jx[c] / eps0);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T dBdx = (By[c] - Byxm[c]) / dx;
            T ez = wall ? T(0.0) : ld32(S.a[S_EZ1 + c], i);
            sEz[c][iz][ix] = ez + dt * (c2 * dBdx - Jz[c] / eps0);
        }
    };

    // own positions: (tx, tz + MZT*k); halo: column ncol and row nrow
    for (int k = 0; k < TM / MZT; ++k) {
        int iz = tz + MZT * k;
        if (tx < ncol && iz < nrow) ampere(tx, iz);
    }
    {
        int h = threadIdx.x;                      // 0 .. MX*MZT-1
        if (h < nrow + 1) ampere(ncol, h);        // halo column (incl. corner)
        else if (h - (nrow + 1) < ncol) ampere(h - (nrow + 1), nrow);   // halo row
    }
    __syncthreads();

    for (int k = 0; k < TM / MZT; ++k) {
        int iz = tz + MZT * k;
        int ix = tx;
        if (ix >= ncol || iz >= nrow) continue;
        const int xr = x0 + ix, zr = z0 + iz;
        const unsigned i = (unsigned)xr + (unsigned)zr * (unsigned)nx;
        // kernel_ym_potential (YM_Physics.cu:589-599)
        T Az[3], Ex[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Az[c] = ld32(S.a[S_AZ1 + c], i);
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
        // sponge / xi_cut of the colour-2/3 fields (YM_Physics.cu:161-210) act
        // right after Faraday on values nothing else reads in between, so they
        // are applied to the stores; the Q2/Q3 part stays in the filter pass.
        const int act = action ? action[xr] : 0;
        const T dec = (act == 1) ? T(decay[xr]) : T(1.0);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T dEx = (sEx[c][iz + 1][ix] - Ex[c]) / dz;
            T dEz = (sEz[c][iz][ixp] - sEz[c][iz][ix]) / dx;
            T by = ld32(S.a[S_BY1 + c], i);
            T byn = by - dt * (dEx - dEz - aYM * cr[c]);
            T exn = Ex[c], ezn = sEz[c][iz][ix], azn = Az[c];
            if (c > 0 && act == 1) { byn *= dec; exn *= dec; ezn *= dec; azn *= dec; }
            if (c > 0 && act == 2) { byn = T(0.0); exn = T(0.0); ezn = T(0.0); azn = T(0.0); }
            st32(O.a[S_BY1 + c], i, byn);
            st32(O.a[S_EX1 + c], i, exn);
            st32(O.a[S_EZ1 + c], i, ezn);
            if (c > 0 || !p.freeze_az1) st32(O.a[S_AZ1 + c], i, azn);
        }
    }
}

// --------------------------------------------------------------------------
// Maxwell, one thread per cell, no LDS: the thread evaluates Ampere at its
// cell (Ex, Ez), Ez at x+1 and Ex at z+1 itself (the neighbours' values come
// from the same formulas on the same inputs, so the results are the bits the
// tiled kernel shares through LDS), then potential and Faraday.  Blocks of
// 64 x 4 cells in XCD-aware row-band order: the z+1 / x+1 re-reads hit L2.
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void jx_at(const State<T>& F, int x, int z, const KParams& p, T Jx[3]) {
    const int nx = p.nx;
    const size_t i = (size_t)x + (size_t)z * nx;
    const size_t ixp = (size_t)xp_src(x, p) + (size_t)z * nx;
    T vxr[2], nQx[2][3];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int b = S_NA + s * F_COUNT;
        const T n0 = F.a[b + F_N][i], np = F.a[b + F_N][ixp];
        vxr[s] = T(0.5) * (F.a[b + F_PX][i] / kmax(n0, T(1e-10)) + F.a[b + F_PX][ixp] / kmax(np, T(1e-10)));
#pragma unroll
        for (int c = 0; c < 3; ++c) nQx[s][c] = T(0.5) * (n0 * F.a[b + F_Q1 + c][i] + np * F.a[b + F_Q1 + c][ixp]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) Jx[c] = -(nQx[0][c] * vxr[0] + nQx[1][c] * vxr[1]);
}

template <typename T>
__device__ __forceinline__ void jz_at(const State<T>& F, int x, int z, const KParams& p, T Jz[3]) {
    const int nx = p.nx;
    const size_t i = (size_t)x + (size_t)z * nx;
    const size_t izp = (size_t)x + (size_t)zp_per(z, p.nz) * nx;
    T vzt[2], nQz[2][3];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int b = S_NA + s * F_COUNT;
        const T n0 = F.a[b + F_N][i], nzp = F.a[b + F_N][izp];
        vzt[s] = T(0.5) * (F.a[b + F_PZ][i] / kmax(n0, T(1e-10)) + F.a[b + F_PZ][izp] / kmax(nzp, T(1e-10)));
#pragma unroll
        for (int c = 0; c < 3; ++c) nQz[s][c] = T(0.5) * (n0 * F.a[b + F_Q1 + c][i] + nzp * F.a[b + F_Q1 + c][izp]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) Jz[c] = -(nQz[0][c] * vzt[0] + nQz[1][c] * vzt[1]);
}

// Ampere (YM_Physics.cu:70-113) for Ex^a at (x, z) from state S and new fluid O.
template <typename T>
__device__ __forceinline__ void ampere_ex(const State<T>& S, const State<T>& O, int x, int z, const KParams& p,
                                          T out[3]) {
    const int nx = p.nx;
    const size_t i = (size_t)x + (size_t)z * nx;
    const size_t izm = (size_t)x + (size_t)zm_per(z, p.nz) * nx;
    T Jx[3];
    jx_at(O, x, z, p, Jx);
    T By[3], Byzm[3], Az[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        By[c] = S.a[S_BY1 + c][i];
        Byzm[c] = S.a[S_BY1 + c][izm];
        Az[c] = S.a[S_AZ1 + c][i];
    }
    const T cr[3] = {Az[1] * By[2] - Az[2] * By[1], Az[2] * By[0] - Az[0] * By[2], Az[0] * By[1] - Az[1] * By[0]};
    const T dt = T(p.dt), c2 = T(p.c2), aYM = T(p.alpha), eps0 = T(p.eps0), dz = T(p.dz);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T dBdz = (By[c] - Byzm[c]) / dz;
        out[c] = S.a[S_EX1 + c][i] + dt * (c2 * (-dBdz - aYM * cr[c]) - Jx[c] / eps0);
    }
}

// Ampere for Ez^a at (x, z); walled x: Ez = 0 at the walls first
// (kernel_enforce_boundaries, YM_Fluid.cu:23-26).
template <typename T>
__device__ __forceinline__ void ampere_ez(const State<T>& S, const State<T>& O, int x, int z, const KParams& p,
                                          T out[3]) {
    const int nx = p.nx;
    const size_t i = (size_t)x + (size_t)z * nx;
    const int xm = p.periodic_x ? (x + nx - 1) % nx : (x == 0 ? 0 : x - 1);
    const size_t ixm = (size_t)xm + (size_t)z * nx;
    T Jz[3];
    jz_at(O, x, z, p, Jz);
    const bool wall = !p.periodic_x && (x == 0 || x == nx - 1);
    const T dt = T(p.dt), c2 = T(p.c2), eps0 = T(p.eps0), dx = T(p.dx);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T dBdx = (S.a[S_BY1 + c][i] - S.a[S_BY1 + c][ixm]) / dx;
        T ez = wall ? T(0.0) : S.a[S_EZ1 + c][i];
        out[c] = ez + dt * (c2 * dBdx - Jz[c] / eps0);
    }
}

constexpr int MC_X = 64, MC_Z = 4;

template <typename T>
__global__ __launch_bounds__(MC_X * MC_Z) void k_maxwell_cell(State<T> S, State<T> O, KParams p, int ntx, int ntz,
                                                              const double* decay, const int* action) {
    const int nx = p.nx, nz = p.nz;
    const int lt = tile_order(xcd_logical(blockIdx.x, gridDim.x), ntx * ntz, p.rev);
    if (lt >= ntx * ntz || lt < 0) return;
    const int x = (lt % ntx) * MC_X + (int)(threadIdx.x % MC_X);
    const int z = (lt / ntx) * MC_Z + (int)(threadIdx.x / MC_X);
    if (x >= nx || z >= nz) return;
    const size_t i = (size_t)x + (size_t)z * nx;
    const T dt = T(p.dt), aYM = T(p.alpha), dx = T(p.dx), dz = T(p.dz);
    T Ex[3], Ez[3], Exzp[3], Ezxp[3];
    ampere_ex(S, O, x, z, p, Ex);
    ampere_ez(S, O, x, z, p, Ez);
    // Faraday's neighbours; walled x: E(x+1) of the last column is its own
    const int xq = (!p.periodic_x && x == nx - 1) ? x : xp_src(x, p);
    if (xq == x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) Ezxp[c] = Ez[c];
    } else {
        ampere_ez(S, O, xq, z, p, Ezxp);
    }
    ampere_ex(S, O, x, zp_per(z, nz), p, Exzp);
    // kernel_ym_potential (YM_Physics.cu:589-599), kernel_ym_faraday (:120-151)
    T Az[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Az[c] = S.a[S_AZ1 + c][i];
        if (c > 0 || !p.freeze_az1) Az[c] = Az[c] - dt * Ez[c];
    }
    const T cr[3] = {Az[1] * Ex[2] - Az[2] * Ex[1], Az[2] * Ex[0] - Az[0] * Ex[2], Az[0] * Ex[1] - Az[1] * Ex[0]};
    const int act = action ? action[x] : 0;
    const T dec = (act == 1) ? T(decay[x]) : T(1.0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T dEx = (Exzp[c] - Ex[c]) / dz;
        T dEz = (Ezxp[c] - Ez[c]) / dx;
        T by = S.a[S_BY1 + c][i];
        T byn = by - dt * (dEx - dEz - aYM * cr[c]);
        T exn = Ex[c], ezn = Ez[c], azn = Az[c];
        if (c > 0 && act == 1) { byn *= dec; exn *= dec; ezn *= dec; azn *= dec; }
        if (c > 0 && act == 2) { byn = T(0.0); exn = T(0.0); ezn = T(0.0); azn = T(0.0); }
        O.a[S_BY1 + c][i] = byn;
        O.a[S_EX1 + c][i] = exn;
        O.a[S_EZ1 + c][i] = ezn;
        if (c > 0 || !p.freeze_az1) O.a[S_AZ1 + c][i] = azn;
    }
}

// --------------------------------------------------------------------------
// Maxwell as a z march.  A wavefront owns 64 adjacent columns and marches down
// a chunk of TZ rows; lane l holds column x = 63*w + l.  Per row it loads the
// new fluid of both species and the step-start fields of its cell once:
//   * Jx at x+1/2 takes the x+1 quantities (px/n, nQ) from lane l+1 by a DPP
//     shift, Jz at z+1/2 the row below from the next iteration;
//   * Ez needs By at x-1 (loaded with the row: the line is lane l-1's, so
//     the load hits the cache), Faraday the new Ez at x+1 (lane l+1) and the
//     new Ex at z+1 (next iteration), so row z is finished one iteration
//     after it is loaded.
// Lane 63 is a halo lane (its x+1 neighbours are in the next wavefront) and
// stores nothing: 63 output columns per wavefront, one extra row per chunk.
// The arithmetic is that of currents_at / ampere / potential / Faraday in
// k_maxwell, operation for operation, so all Maxwell variants agree bit for
// bit.  PF = 1 loads row e+1 before the arithmetic of row e.
// --------------------------------------------------------------------------
constexpr int MZM_OUT = 63;

template <typename T>
struct MRow {                  // one row of one column: inputs of the march
    T n[2], px[2], pz[2], q[2][3];
    T by[3], byxm[3], az[3], ex[3], ez[3];
};

template <typename T>
__device__ __forceinline__ void mrow_load(const State<T>& S, const State<T>& O, unsigned i, unsigned ixm,
                                          MRow<T>& r) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int b = S_NA + s * F_COUNT;
        r.n[s] = ld32(O.a[b + F_N], i);
        r.px[s] = ld32(O.a[b + F_PX], i);
        r.pz[s] = ld32(O.a[b + F_PZ], i);
#pragma unroll
        for (int c = 0; c < 3; ++c) r.q[s][c] = ld32(O.a[b + F_Q1 + c], i);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r.by[c] = ld32(S.a[S_BY1 + c], i);
        r.byxm[c] = ld32(S.a[S_BY1 + c], ixm);
        r.az[c] = ld32(S.a[S_AZ1 + c], i);
        r.ex[c] = ld32(S.a[S_EX1 + c], i);
        r.ez[c] = ld32(S.a[S_EZ1 + c], i);
    }
}

// One wavefront of the march: wave wl (uniform) of a grid of nwx column
// groups x chunks of TZ rows.
template <typename T, int TZ, int PF, bool EN = false>
__device__ __forceinline__ void maxwell_zm_wave(const State<T>& S, const State<T>& O, const KParams& p, int nwx,
                                                int wl, int lane, const double* decay, const int* action,
                                                double* epart = nullptr) {
    const int nx = p.nx, nz = p.nz;
    const bool per = p.periodic_x != 0;
    const int x = (wl % nwx) * MZM_OUT + lane;
    // lanes past the last column mirror the x+1 neighbour of column nx-1
    const int xc = x < nx ? x : (per ? x % nx : nx - 1);
    const int xm = per ? (xc + nx - 1) % nx : (xc == 0 ? 0 : xc - 1);
    const bool wall = !per && (xc == 0 || xc == nx - 1);     // YM_Fluid.cu:23-26
    const bool last_wall = !per && xc == nx - 1;              // Faraday: E(x+1) := E(x)
    const bool store = lane < MZM_OUT && x < nx;
    const int z0 = (wl / nwx) * TZ;
    const int z1 = min(z0 + TZ, nz);
    const T dt = T(p.dt), c2 = T(p.c2), aYM = T(p.alpha);
    const CDiv<T> ddx = cdiv_dx<T>(p), ddz = cdiv_dz<T>(p), deps0 = cdiv_eps0<T>(p);
    const int act = action ? action[xc] : 0;
    const T dec = (act == 1) ? T(decay[xc]) : T(1.0);
    // halo lane 63 and lanes past the last column store through range-checked
    // buffer stores at an out-of-range offset: no branch around the stores
    const unsigned bytes = (unsigned)nx * (unsigned)nz * (unsigned)sizeof(T);
    __amdgpu_buffer_rsrc_t rby[3], rex[3], rez[3], raz[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        rby[c] = buf_rsrc(O.a[S_BY1 + c], bytes);
        rex[c] = buf_rsrc(O.a[S_EX1 + c], bytes);
        rez[c] = buf_rsrc(O.a[S_EZ1 + c], bytes);
        raz[c] = buf_rsrc(O.a[S_AZ1 + c], bytes);
    }

    // carried from row e-1
    T byp[3], byxmp[3], azp[3], ezp[3], exn_p[3];
    T uzp[2], nqp[2][3];
    {
        const unsigned im = (unsigned)xc + (unsigned)zm_per(z0, nz) * (unsigned)nx;
#pragma unroll
        for (int c = 0; c < 3; ++c) byp[c] = ld32(S.a[S_BY1 + c], im);
    }
    MRow<T> cur, nxt;
    // EN: the energy density of the wave's own cells after this step
    // (YM_Energy.cu:3-21: kinetic 0.5 |p|^2 / max(n, 1e-10) of both species,
    // 0.5 (E^2 + By^2) of the filtered fields), a compensated sum per lane
    CSum eacc{0.0, 0.0};
    auto rbase = [&](int e) { return (unsigned)(e < nz ? e : e - nz) * (unsigned)nx; };
    mrow_load(S, O, rbase(z0) + (unsigned)xc, rbase(z0) + (unsigned)xm, cur);

#pragma unroll 1
    for (int e = z0; e <= z1; ++e) {
        if (PF && e < z1) mrow_load(S, O, rbase(e + 1) + (unsigned)xc, rbase(e + 1) + (unsigned)xm, nxt);
        // cell quantities of the currents (kernel_ym_currents, YM_Physics.cu:9-60)
        T ux[2], uz[2], nq[2][3];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const T nf = kmax(cur.n[s], T(1e-10));
            ux[s] = cur.px[s] / nf;
            uz[s] = cur.pz[s] / nf;
#pragma unroll
            for (int c = 0; c < 3; ++c) nq[s][c] = cur.n[s] * cur.q[s][c];
        }
        if constexpr (EN) {
            if (store && e < z1) {
                double ke = 0.0;
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const double nn = kmax((double)cur.n[s], 1e-10);
                    const double px = cur.px[s], pz = cur.pz[s];
                    ke = ke + 0.5 * (px * px + pz * pz) / nn;
                }
                cadd(eacc, ke);
            }
        }
        // Jx at (x+1/2, e) with the x+1 values of lane l+1
        T Jx[3];
        {
            T vxr[2], nQx[2][3];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                vxr[s] = T(0.5) * (ux[s] + lane_from_right(ux[s]));
#pragma unroll
                for (int c = 0; c < 3; ++c) nQx[s][c] = T(0.5) * (nq[s][c] + lane_from_right(nq[s][c]));
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) Jx[c] = -(nQx[0][c] * vxr[0] + nQx[1][c] * vxr[1]);
        }
        // Ampere for Ex at row e (YM_Physics.cu:70-113)
        T exn[3];
        {
            const T cr[3] = {cur.az[1] * cur.by[2] - cur.az[2] * cur.by[1],
                             cur.az[2] * cur.by[0] - cur.az[0] * cur.by[2],
                             cur.az[0] * cur.by[1] - cur.az[1] * cur.by[0]};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                T dBdz = ddz(cur.by[c] - byp[c]);
                exn[c] = cur.ex[c] + dt * (c2 * (-dBdz - aYM * cr[c]) - deps0(Jx[c]));
            }
        }
        if (e > z0) {
            // row e-1: Jz at (x, e-1/2), Ampere for Ez, potential, Faraday
            T ezn[3];
            {
                T vzt[2], nQz[2][3];
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    vzt[s] = T(0.5) * (uzp[s] + uz[s]);
#pragma unroll
                    for (int c = 0; c < 3; ++c) nQz[s][c] = T(0.5) * (nqp[s][c] + nq[s][c]);
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    T Jz = -(nQz[0][c] * vzt[0] + nQz[1][c] * vzt[1]);
                    T dBdx = ddx(byp[c] - byxmp[c]);
                    T ez = wall ? T(0.0) : ezp[c];
                    ezn[c] = ez + dt * (c2 * dBdx - deps0(Jz));
                }
            }
            T ezr[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                ezr[c] = lane_from_right(ezn[c]);
                if (last_wall) ezr[c] = ezn[c];
            }
            // kernel_ym_potential (YM_Physics.cu:589-599), kernel_ym_faraday (:120-151)
            T Az[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                Az[c] = azp[c];
                if (c > 0 || !p.freeze_az1) Az[c] = Az[c] - dt * ezn[c];
            }
            const T cr[3] = {Az[1] * exn_p[2] - Az[2] * exn_p[1],
                             Az[2] * exn_p[0] - Az[0] * exn_p[2],
                             Az[0] * exn_p[1] - Az[1] * exn_p[0]};
            {
                const unsigned off = store ? (rbase(e - 1) + (unsigned)xc) * (unsigned)sizeof(T) : kNoStore;
                double em = 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    T dEx = ddz(exn[c] - exn_p[c]);
                    T dEz = ddx(ezr[c] - ezn[c]);
                    T byn = byp[c] - dt * (dEx - dEz - aYM * cr[c]);
                    T exo = exn_p[c], ezo = ezn[c], azo = Az[c];
                    if (c > 0 && act == 1) { byn *= dec; exo *= dec; ezo *= dec; azo *= dec; }
                    if (c > 0 && act == 2) { byn = T(0.0); exo = T(0.0); ezo = T(0.0); azo = T(0.0); }
                    buf_st(byn, rby[c], off);
                    buf_st(exo, rex[c], off);
                    buf_st(ezo, rez[c], off);
                    if (c > 0 || !p.freeze_az1) buf_st(azo, raz[c], off);
                    if constexpr (EN) {
                        const double a = exo, b = ezo, f = byn;
                        em = em + a * a + b * b + f * f;
                    }
                }
                if constexpr (EN) {
                    if (store) cadd(eacc, 0.5 * em);
                }
            }
        }
        // carry row e
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            byp[c] = cur.by[c]; byxmp[c] = cur.byxm[c]; azp[c] = cur.az[c]; ezp[c] = cur.ez[c]; exn_p[c] = exn[c];
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            uzp[s] = uz[s];
#pragma unroll
            for (int c = 0; c < 3; ++c) nqp[s][c] = nq[s][c];
        }
        if (e < z1) {
            if (PF) cur = nxt;
            else mrow_load(S, O, rbase(e + 1) + (unsigned)xc, rbase(e + 1) + (unsigned)xm, cur);
        }
    }
    if constexpr (EN) {
        // the wave's lanes merged in a fixed order; one CSum per wave
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const CSum o{__shfl_down(eacc.s, off, 64), __shfl_down(eacc.c, off, 64)};
            eacc = cmerge(eacc, o);
        }
        if (lane == 0) {
            epart[2 * (size_t)wl] = eacc.s;
            epart[2 * (size_t)wl + 1] = eacc.c;
        }
    }
}

template <typename T, int TZ, int PF, bool EN = false>
__global__ __launch_bounds__(256) void k_maxwell_zm(State<T> S, State<T> O, KParams p, int nwx, int nch,
                                                   const double* decay, const int* action, double* epart) {
    const int lane = threadIdx.x & 63;
    // the wavefront index is uniform: scalar loop control
    int wl = __builtin_amdgcn_readfirstlane(xcd_logical(blockIdx.x, gridDim.x) * 4 + (int)(threadIdx.x >> 6));
    if (wl >= nwx * nch) return;             // whole wavefronts leave together
    wl = tile_order(wl, nwx * nch, p.rev);
    maxwell_zm_wave<T, TZ, PF, EN>(S, O, p, nwx, wl, lane, decay, action, epart);
}

// --------------------------------------------------------------------------
// Fused z sweep + Maxwell (YM).  A workgroup is three wavefronts on the same
// 64 columns (lane l: column x = 63*w + l, as in k_maxwell_zm) and z chunk.
// Wavefronts 0 and 1 march species 0 and 1 through the z-sweep pipeline of
// k_fct_z; for every finished row they store the new fluid and hand it to
// wavefront 2 through LDS (double-buffered, one barrier per row), which runs
// the Maxwell march step of k_maxwell_zm on both species' new fluid.  The new
// fluid is thus never read back from memory, and the fields are read once
// for the z sources and Maxwell together (Maxwell's read of a row a few
// iterations after the z sources' hits the cache).  The pipelines run one row
// past the chunk, because the Maxwell step of the chunk's last row needs the
// new fluid of the row below; in the last chunk that row is row 0 (z is
// periodic for Maxwell but clamped for the FCT), whose new fluid a second,
// 7-row pipeline computes from the clamped rows 0,0,0,0,1,2,3 exactly as the
// chunk that owns it does.  Every value is computed by the expressions of
// k_fct_z and k_maxwell_zm in the same order, so the results are those of the
// split kernels bit for bit.
// --------------------------------------------------------------------------
template <typename T, int TZ>
__global__ __launch_bounds__(192) void k_fct_zmx(State<T> S, Stage<T> X, State<T> O, KParams p, int nwx, int nch,
                                                const double* decay, const int* action) {
    typedef YMModel M;
    constexpr int NV = M::NV;
    typedef typename M::template Row<T> Row;
    __shared__ T sF[2][2][NV][64];             // [buffer][species][variable][lane]
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;           // 0, 1: z sweep of that species; 2: Maxwell
    int lb = xcd_logical(blockIdx.x, gridDim.x);
    if (lb >= nwx * nch) return;               // whole workgroups leave together
    lb = tile_order(lb, nwx * nch, p.rev);
    const int nx = p.nx, nz = p.nz;
    const bool per = p.periodic_x != 0;
    const int x = (lb % nwx) * MZM_OUT + lane;
    const int xc = x < nx ? x : (per ? x % nx : nx - 1);
    const int xm = per ? (xc + nx - 1) % nx : (xc == 0 ? 0 : xc - 1);      // xm_src(xc)
    const bool wall = !per && (xc == 0 || xc == nx - 1);
    const bool store = lane < MZM_OUT && x < nx;
    const int z0 = (lb / nwx) * TZ;
    const int zend = min(z0 + TZ, nz);
    const bool last = zend == nz;
    const int r_end = last ? nz : zend;        // Maxwell steps z0 .. r_end (r_end: the row below)
    const T dt = T(p.dt);
    auto rbase = [&](int r) { return (unsigned)(r < nz ? r : r - nz) * (unsigned)nx; };

    if (wv == 2) {
        // ---- Maxwell march (k_maxwell_zm) on the new fluid from LDS ----
        const bool last_wall = !per && xc == nx - 1;      // Faraday: E(x+1) := E(x)
        const T c2 = T(p.c2), aYM = T(p.alpha), eps0 = T(p.eps0), dx = T(p.dx), dz = T(p.dz);
        const int act = action ? action[xc] : 0;
        const T dec = (act == 1) ? T(decay[xc]) : T(1.0);
        T byp[3], byxmp[3], azp[3], ezp[3], exn_p[3], uzp[2], nqp[2][3];
        T fby[3], fbyxm[3], faz[3], fex[3], fez[3];       // field row of the next step
        auto fload = [&](int r) {
            const unsigned i = rbase(r) + (unsigned)xc, im = rbase(r) + (unsigned)xm;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                fby[c] = ld32(S.a[S_BY1 + c], i);
                fbyxm[c] = ld32(S.a[S_BY1 + c], im);
                faz[c] = ld32(S.a[S_AZ1 + c], i);
                fex[c] = ld32(S.a[S_EX1 + c], i);
                fez[c] = ld32(S.a[S_EZ1 + c], i);
            }
        };
        {
            const unsigned im = (unsigned)zm_per(z0, nz) * (unsigned)nx + (unsigned)xc;
#pragma unroll
            for (int c = 0; c < 3; ++c) byp[c] = ld32(S.a[S_BY1 + c], im);
        }
        fload(z0);
        int buf = 0;
#pragma unroll 1
        for (int r = z0; r <= r_end; ++r) {
            T by[3], byxm[3], az[3], ex[3], ez[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) { by[c] = fby[c]; byxm[c] = fbyxm[c]; az[c] = faz[c]; ex[c] = fex[c]; ez[c] = fez[c]; }
            if (r < r_end) fload(r + 1);
            __syncthreads();                       // the new fluid of row r is in sF[buf]
            T ux[2], uz[2], nq[2][3];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const T n = sF[buf][s][F_N][lane];
                const T nf = kmax(n, T(1e-10));
                ux[s] = sF[buf][s][F_PX][lane] / nf;
                uz[s] = sF[buf][s][F_PZ][lane] / nf;
#pragma unroll
                for (int c = 0; c < 3; ++c) nq[s][c] = n * sF[buf][s][F_Q1 + c][lane];
            }
            buf ^= 1;
            // Jx at (x+1/2, r) with the x+1 values of lane l+1 (kernel_ym_currents)
            T Jx[3];
            {
                T vxr[2], nQx[2][3];
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    vxr[s] = T(0.5) * (ux[s] + lane_from_right(ux[s]));
#pragma unroll
                    for (int c = 0; c < 3; ++c) nQx[s][c] = T(0.5) * (nq[s][c] + lane_from_right(nq[s][c]));
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) Jx[c] = -(nQx[0][c] * vxr[0] + nQx[1][c] * vxr[1]);
            }
            // Ampere for Ex at row r
            T exn[3];
            {
                const T cr[3] = {az[1] * by[2] - az[2] * by[1], az[2] * by[0] - az[0] * by[2],
                                 az[0] * by[1] - az[1] * by[0]};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    T dBdz = (by[c] - byp[c]) / dz;
                    exn[c] = ex[c] + dt * (c2 * (-dBdz - aYM * cr[c]) - Jx[c] / eps0);
                }
            }
            if (r > z0) {
                // row r-1: Jz, Ampere for Ez, potential, Faraday
                T ezn[3];
                {
                    T vzt[2], nQz[2][3];
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        vzt[s] = T(0.5) * (uzp[s] + uz[s]);
#pragma unroll
                        for (int c = 0; c < 3; ++c) nQz[s][c] = T(0.5) * (nqp[s][c] + nq[s][c]);
                    }
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        T Jz = -(nQz[0][c] * vzt[0] + nQz[1][c] * vzt[1]);
                        T dBdx = (byp[c] - byxmp[c]) / dx;
                        T ez0 = wall ? T(0.0) : ezp[c];
                        ezn[c] = ez0 + dt * (c2 * dBdx - Jz / eps0);
                    }
                }
                T ezr[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    ezr[c] = lane_from_right(ezn[c]);
                    if (last_wall) ezr[c] = ezn[c];
                }
                T Az[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    Az[c] = azp[c];
                    if (c > 0 || !p.freeze_az1) Az[c] = Az[c] - dt * ezn[c];
                }
                const T cr[3] = {Az[1] * exn_p[2] - Az[2] * exn_p[1], Az[2] * exn_p[0] - Az[0] * exn_p[2],
                                 Az[0] * exn_p[1] - Az[1] * exn_p[0]};
                if (store) {
                    const unsigned i = rbase(r - 1) + (unsigned)xc;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        T dEx = (exn[c] - exn_p[c]) / dz;
                        T dEz = (ezr[c] - ezn[c]) / dx;
                        T byn = byp[c] - dt * (dEx - dEz - aYM * cr[c]);
                        T exo = exn_p[c], ezo = ezn[c], azo = Az[c];
                        if (c > 0 && act == 1) { byn *= dec; exo *= dec; ezo *= dec; azo *= dec; }
                        if (c > 0 && act == 2) { byn = T(0.0); exo = T(0.0); ezo = T(0.0); azo = T(0.0); }
                        st32(O.a[S_BY1 + c], i, byn);
                        st32(O.a[S_EX1 + c], i, exo);
                        st32(O.a[S_EZ1 + c], i, ezo);
                        if (c > 0 || !p.freeze_az1) st32(O.a[S_AZ1 + c], i, azo);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) { byp[c] = by[c]; byxmp[c] = byxm[c]; azp[c] = az[c]; ezp[c] = ez[c]; exn_p[c] = exn[c]; }
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                uzp[s] = uz[s];
#pragma unroll
                for (int c = 0; c < 3; ++c) nqp[s][c] = nq[s][c];
            }
        }
        return;
    }

    // ---- z-sweep pipeline of species wv (k_fct_z with next-row prefetch) ----
    const int sp = wv;
    const int b = M::base(sp);
    const T dtdz = T(p.dtdz);
    auto zload = [&](int r, Row& row, T (&F)[NV], T (&V)[NV]) {
        const unsigned i = (unsigned)xc + (unsigned)r * (unsigned)nx;
        M::load_row(S, i, (unsigned)xm + (unsigned)r * (unsigned)nx, row);
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v] = ld32(S.a[b + v], i);
#pragma unroll
        for (int v = 0; v < NV; ++v) V[v] = ld32(X.a[sp * NV + v], i);
    };
    int buf = 0;
    // the chunk's own rows z0 .. (last ? nz-1 : zend) -- zend only for the
    // Maxwell step -- then, in the last chunk, row 0 again for the step of
    // row nz-1
#pragma unroll 1
    for (int pass = 0; pass < (last ? 2 : 1); ++pass) {
        const bool own = pass == 0;
        const int q0 = own ? z0 : 0;
        const int elast = own ? (last ? nz + 2 : zend + 3) : 3;
        int rc = clampi(q0 - 3, 0, nz - 1);
        Row prev, cur;
        T F[NV], V[NV];
        {
            T Fd[NV], Vd[NV];
            zload(zm_per(rc, nz), prev, Fd, Vd);
        }
        zload(rc, cur, F, V);
        ZPipe<T, NV> zp;
        zp.Um = T(0);
        zp.MUm = T(0);
#pragma unroll 1
        for (int e = q0 - 3; e <= elast; ++e) {
            const int rn = clampi(min(e + 1, elast), 0, nz - 1);
            Row nrow;
            T nF[NV], nV[NV];
            zload(rn, nrow, nF, nV);
            T u, Sv[NV], out[NV];
            M::z_src_rows(F, cur, prev, S, sp, xc, rc, p, u, Sv);
            if (zp.push(e, q0, V, u, Sv, dt, dtdz, out)) {
                const int r = e - 3;
                if (out[0] < T(0.05)) out[0] = T(0.05);          // kernel_enforce_boundaries
                if (wall) out[M::IPX] = T(0.0);
                if (own && store && r < zend) {
                    const unsigned io = (unsigned)xc + (unsigned)r * (unsigned)nx;
#pragma unroll
                    for (int v = 0; v < NV; ++v) st32(O.a[b + v], io, out[v]);
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) sF[buf][sp][v][lane] = out[v];
                __syncthreads();
                buf ^= 1;
            }
            M::sel_row(rn != rc, cur, prev);
            cur = nrow;
#pragma unroll
            for (int v = 0; v < NV; ++v) { F[v] = nF[v]; V[v] = nV[v]; }
            rc = rn;
        }
    }
}

// ==========================================================================
// EMHD Maxwell: enforce_boundaries (Ey, Ez, Bx walls) + currents + Ampere +
// vector potential + Faraday (emhdgpu2d/main_coupled.cu:48-63, 99-110;
// EMHD_Physics.cu:21-98).  Same tiling as k_maxwell; the updated E of the
// tile plus one halo column on each side and one halo row is kept in LDS
// (the walled Faraday at x = nx-1 uses E(x-1)).
// ==========================================================================
template <typename T>
__global__ __launch_bounds__(MX * MZT) void k_emhd_maxwell(State<T> S, State<T> O, KParams p, int ntx, int ntz) {
    __shared__ T sE[3][MZ + 1][MX + 2];     // column index ix + 1
    const int nx = p.nx, nz = p.nz;
    const int lt = tile_order(xcd_logical(blockIdx.x, gridDim.x), ntx * ntz, p.rev);
    if (lt >= ntx * ntz || lt < 0) return;
    const int x0 = (lt % ntx) * MX, z0 = (lt / ntx) * MZ;
    const int ncol = min(MX, nx - x0), nrow = min(MZ, nz - z0);
    const T dt = T(p.dt), c2 = T(p.c2), eps0 = T(p.eps0);
    const T dx = T(p.dx), dz = T(p.dz);
    const bool per = p.periodic_x != 0;
    const int tx = threadIdx.x % MX, tz = threadIdx.x / MX;
    auto wall = [&](int x) { return !per && (x == 0 || x == nx - 1); };

    auto ampere = [&](int ix, int iz) {
        int xr;
        if (ix < 0) xr = x0 - 1;
        else if (ix < ncol) xr = x0 + ix;
        else xr = xp_src(x0 + ncol - 1, p);
        const int zr = (iz < nrow) ? z0 + iz : zp_per(z0 + nrow - 1, nz);
        const size_t i = (size_t)xr + (size_t)zr * nx;
        const int zm = zm_per(zr, nz);
        const size_t izm = (size_t)xr + (size_t)zm * nx;
        // currents (main_coupled.cu:99-110), summed over the fluids
        const int xp = xp_src(xr, p), zp = zp_per(zr, nz);
        const size_t ixp = (size_t)xp + (size_t)zr * nx, izp = (size_t)xr + (size_t)zp * nx;
        T jx = T(0), jy = T(0), jz = T(0);
        for (int s = 0; s < p.nsp; ++s) {
            const int b = em_base(s);
            const T q = T(s ? p.q1 : p.q0);
            const T h = T((s ? p.q1 : p.q0) * 0.5);
            const T a = h * (O.a[b + EV_PX][i] + O.a[b + EV_PX][ixp]);
            const T bb = q * O.a[b + EV_PY][i];
            const T c = h * (O.a[b + EV_PZ][i] + O.a[b + EV_PZ][izp]);
            if (s == 0) { jx = a; jy = bb; jz = c; }
            else { jx = jx + a; jy = jy + bb; jz = jz + c; }
        }
        const bool w = wall(xr);
        const T Bx0 = w ? T(0) : S.a[E_BX][i], Bxzm = w ? T(0) : S.a[E_BX][izm];
        T dBy_dz = (S.a[E_BY][i] - S.a[E_BY][izm]) / dz;
        T curlB_x = -dBy_dz;
        T dBx_dz = (Bx0 - Bxzm) / dz;
        T dBz_dx, dBy_dx;
        if (per) {
            const size_t ixm = (size_t)((xr + nx - 1) % nx) + (size_t)zr * nx;
            dBz_dx = (S.a[E_BZ][i] - S.a[E_BZ][ixm]) / dx;
            dBy_dx = (S.a[E_BY][i] - S.a[E_BY][ixm]) / dx;
        } else if (xr == 0) {
            const size_t i1 = (size_t)1 + (size_t)zr * nx;
            dBz_dx = (S.a[E_BZ][i1] - S.a[E_BZ][i]) / dx;
            dBy_dx = (S.a[E_BY][i1] - S.a[E_BY][i]) / dx;
        } else {
            const size_t ixm = (size_t)(xr - 1) + (size_t)zr * nx;
            dBz_dx = (S.a[E_BZ][i] - S.a[E_BZ][ixm]) / dx;
            dBy_dx = (S.a[E_BY][i] - S.a[E_BY][ixm]) / dx;
        }
        T curlB_y = dBx_dz - dBz_dx;
        T curlB_z = dBy_dx;
        const T ex = S.a[E_EX][i];
        const T ey = w ? T(0) : S.a[E_EY][i];
        const T ez = w ? T(0) : S.a[E_EZ][i];
        sE[0][iz][ix + 1] = ex + dt * (c2 * curlB_x - jx / eps0);
        sE[1][iz][ix + 1] = ey + dt * (c2 * curlB_y - jy / eps0);
        sE[2][iz][ix + 1] = ez + dt * (c2 * curlB_z - jz / eps0);
    };

    for (int k = 0; k < MZ / MZT; ++k) {
        int iz = tz + MZT * k;
        if (tx < ncol && iz < nrow) ampere(tx, iz);
    }
    {
        int h = threadIdx.x;
        if (h < nrow + 1) ampere(ncol, h);                         // right halo column (+ corner)
        else if (h - (nrow + 1) < ncol) ampere(h - (nrow + 1), nrow);   // halo row
        else if (x0 > 0 && h - (nrow + 1) - ncol < nrow) ampere(-1, h - (nrow + 1) - ncol);  // left halo column
    }
    __syncthreads();

    for (int k = 0; k < MZ / MZT; ++k) {
        const int iz = tz + MZT * k, ix = tx;
        if (ix >= ncol || iz >= nrow) continue;
        const int xr = x0 + ix, zr = z0 + iz;
        const size_t i = (size_t)xr + (size_t)zr * nx;
        const T Ex = sE[0][iz][ix + 1], Ey = sE[1][iz][ix + 1], Ez = sE[2][iz][ix + 1];
        T dEy_dz = (sE[1][iz + 1][ix + 1] - Ey) / dz;
        T curlE_x = -dEy_dz;
        T dEx_dz = (sE[0][iz + 1][ix + 1] - Ex) / dz;
        T dEz_dx, dEy_dx;
        if (!per && xr == nx - 1) {
            dEz_dx = (Ez - sE[2][iz][ix]) / dx;
            dEy_dx = (Ey - sE[1][iz][ix]) / dx;
        } else {
            dEz_dx = (sE[2][iz][ix + 2] - Ez) / dx;
            dEy_dx = (sE[1][iz][ix + 2] - Ey) / dx;
        }
        T curlE_y = dEx_dz - dEz_dx;
        T curlE_z = dEy_dx;
        const bool w = wall(xr);
        const T bx = w ? T(0) : S.a[E_BX][i];
        O.a[E_BX][i] = bx - dt * curlE_x;
        O.a[E_BY][i] = S.a[E_BY][i] - dt * curlE_y;
        O.a[E_BZ][i] = S.a[E_BZ][i] - dt * curlE_z;
        O.a[E_EX][i] = Ex;
        O.a[E_EY][i] = Ey;
        O.a[E_EZ][i] = Ez;
        O.a[E_AX][i] = S.a[E_AX][i] - dt * Ex;
        O.a[E_AY][i] = S.a[E_AY][i] - dt * Ey;
        O.a[E_AZ][i] = S.a[E_AZ][i] - dt * Ez;
    }
}

// ==========================================================================
// launchers
// ==========================================================================
template <typename T, int XB, typename M>
static void launch_fct_x_t(const State<T>& S, const Stage<T>& out, const KParams& p, int sp0, int nsp,
                           int rows, int vs, hipStream_t st) {
    const int ntx = (p.nx + XB - 1) / XB;
    const int nrg = (p.nz + rows - 1) / rows;
    dim3 grid(round8(ntx * nrg * nsp));
    if (vs)
        hipLaunchKernelGGL((k_fct_x<T, XB, M, 1>), grid, dim3(XB), 0, st, S, out, p, sp0, nsp, ntx, rows);
    else
        hipLaunchKernelGGL((k_fct_x<T, XB, M, 0>), grid, dim3(XB), 0, st, S, out, p, sp0, nsp, ntx, rows);
}

template <typename T, int C>
static void launch_fct_xr_t(const State<T>& S, const Stage<T>& out, const KParams& p, int sp0, int nsp,
                            hipStream_t st) {
    const int nseg = (p.nx + XR_OUT_LANES * C - 1) / (XR_OUT_LANES * C);
    const int waves = nsp * nseg * p.nz;
    hipLaunchKernelGGL((k_fct_xr<T, C>), dim3(round8((waves + 3) / 4)), dim3(256), 0, st, S, out, p, sp0, nsp,
                       nseg);
}

// Variant v < 18: LDS-staged segments of 64 / 128 / 256 cells (v % 3), rows
// per workgroup 1 / 4 / 16 ((v / 3) % 3), all variables / one variable at a
// time (v / 9).  v = 18 / 19: register form with 4 / 8 cells per lane (YM
// model; the EMHD model takes the default LDS shape instead).
template <typename T, typename M>
void launch_fct_x(const State<T>& S, const Stage<T>& out, const KParams& p, int sp0, int nsp,
                  int variant, hipStream_t st) {
    if (variant >= 18) {
        if constexpr (std::is_same<M, YMModel>::value) {
            if (variant == 18) launch_fct_xr_t<T, 4>(S, out, p, sp0, nsp, st);
            else launch_fct_xr_t<T, 8>(S, out, p, sp0, nsp, st);
            return;
        }
        variant = 14;
    }
    static const int kRows[3] = {1, 4, 16};
    const int rows = kRows[(variant / 3) % 3];
    const int vs = variant / 9;
    switch (variant % 3) {
        case 0: launch_fct_x_t<T, 64, M>(S, out, p, sp0, nsp, rows, vs, st); break;
        case 1: launch_fct_x_t<T, 128, M>(S, out, p, sp0, nsp, rows, vs, st); break;
        default: launch_fct_x_t<T, 256, M>(S, out, p, sp0, nsp, rows, vs, st); break;
    }
}

template <typename T, int TZ, int PF, typename M>
static void launch_fct_z_t(const State<T>& S, const Stage<T>& X, const State<T>& O, const KParams& p,
                           int sp0, int nsp, int grouped, hipStream_t st) {
    const int nchunk = (p.nz + TZ - 1) / TZ;
    const int ntx = (p.nx + ZB_X - 1) / ZB_X;
    if (PF == 2) {
        const int ncg = (nchunk + ZB_C - 1) / ZB_C;
        hipLaunchKernelGGL((k_fct_z2<T, TZ, M>), dim3(round8(nsp * ntx * ncg)), dim3(ZB_X * ZB_C), 0, st, S, X, O, p,
                           sp0, nsp, ntx, ncg);
        return;
    }
    if (grouped) {
        const int ncg = (nchunk + 1) / 2;
        hipLaunchKernelGGL((k_fct_zg<T, TZ, PF, M>), dim3(round8(nsp * ntx * ncg)), dim3(256), 0, st, S, X, O, p,
                           sp0, nsp, ntx, ncg);
        return;
    }
    const int ncg = (nchunk + ZB_C - 1) / ZB_C;
    dim3 grid(round8(nsp * ntx * ncg));
    hipLaunchKernelGGL((k_fct_z<T, TZ, PF, M>), grid, dim3(ZB_X * ZB_C), 0, st, S, X, O, p, sp0, nsp, ntx, ncg);
}

// Variant v: chunk height 16 / 32 / 64 rows (v % 3; halo rows re-read per
// chunk vs. independent column marches in flight), prefetch of the next row
// off / on ((v / 3) % 2), all variables on one lane / split into two groups
// on two waves (v / 6).
template <typename T, typename M>
void launch_fct_z(const State<T>& S, const Stage<T>& X, const State<T>& O, const KParams& p,
                  int sp0, int nsp, int variant, hipStream_t st) {
    switch (variant) {                       // two-row prefetch
        case 12: launch_fct_z_t<T, 16, 2, M>(S, X, O, p, sp0, nsp, 0, st); return;
        case 13: launch_fct_z_t<T, 32, 2, M>(S, X, O, p, sp0, nsp, 0, st); return;
        case 14: launch_fct_z_t<T, 64, 2, M>(S, X, O, p, sp0, nsp, 0, st); return;
        default: break;
    }
    const int gr = variant / 6;
    switch (variant % 6) {
        case 0: launch_fct_z_t<T, 16, 0, M>(S, X, O, p, sp0, nsp, gr, st); break;
        case 1: launch_fct_z_t<T, 32, 0, M>(S, X, O, p, sp0, nsp, gr, st); break;
        case 2: launch_fct_z_t<T, 64, 0, M>(S, X, O, p, sp0, nsp, gr, st); break;
        case 3: launch_fct_z_t<T, 16, 1, M>(S, X, O, p, sp0, nsp, gr, st); break;
        case 5: launch_fct_z_t<T, 64, 1, M>(S, X, O, p, sp0, nsp, gr, st); break;
        default: launch_fct_z_t<T, 32, 1, M>(S, X, O, p, sp0, nsp, gr, st); break;
    }
}

// --------------------------------------------------------------------------
// EMHD Maxwell as a z march (the k_emhd_maxwell update, same expressions).
// Lane l of a wavefront owns column 62 g - 1 + l; lanes 0 and 63 are halo
// lanes, so a wavefront stores the 62 columns of lanes 1..62.  Ampere for
// row e needs the new pz of row e+1 (Jz) and the old Bx, By of row e-1 (the
// previous iteration's row); the x+1 momentum of Jx comes from lane l+1, B
// at the Ampere x-neighbour (x-1, or x+1 at the x = 0 wall) is loaded through
// the reference's index rule.  Faraday for row e-1 runs one iteration later,
// when the new E of row e exists; the new Ey, Ez of x+1 (or x-1 at the
// x = nx-1 wall) come from lane l+1 (l-1).  Lane 63's new Ex lacks its x+1
// momentum but no stored lane reads it; lane 0's and 63's new Ey, Ez are
// complete.  A chunk of TZ rows computes Ampere for TZ+1 rows (the last is
// the next chunk's first, z periodic).
// --------------------------------------------------------------------------
constexpr int EMZ_OUT = 62;

template <typename T>
struct EmFieldRow {
    T ex, ey, ez, bx, by, bz, byo, bzo;   // old fields at (x, z); By, Bz at Ampere's x-neighbour
};

template <typename T, int TZ>
__global__ __launch_bounds__(256) void k_emhd_maxwell_zm(State<T> S, State<T> O, KParams p, int nwx, int nch) {
    const int lane = threadIdx.x & 63;
    int wl = __builtin_amdgcn_readfirstlane(xcd_logical(blockIdx.x, gridDim.x) * 4 + (int)(threadIdx.x >> 6));
    if (wl >= nwx * nch) return;             // whole wavefronts leave together
    wl = tile_order(wl, nwx * nch, p.rev);
    const int nx = p.nx, nz = p.nz;
    const bool per = p.periodic_x != 0;
    const int x = (wl % nwx) * EMZ_OUT - 1 + lane;
    const int xc = map_x(x, p);
    const bool store = lane >= 1 && lane <= EMZ_OUT && x < nx;
    const bool w = !per && (xc == 0 || xc == nx - 1);
    const bool fwd = !per && xc == 0;                           // Ampere: B(1) - B(0) at the x = 0 wall
    const int xo = per ? (xc + nx - 1) % nx : (xc == 0 ? 1 : xc - 1);
    const bool bwd = !per && xc == nx - 1;                      // Faraday: E(x) - E(x-1) at the x = nx-1 wall
    const int z0 = (wl / nwx) * TZ;
    const int z1 = min(z0 + TZ, nz);
    const int nsp = p.nsp;
    const T dt = T(p.dt), c2 = T(p.c2), eps0 = T(p.eps0);
    const T dx = T(p.dx), dz = T(p.dz);
    const unsigned bytes = (unsigned)nx * (unsigned)nz * (unsigned)sizeof(T);
    __amdgpu_buffer_rsrc_t ro[9];
#pragma unroll
    for (int a = 0; a < 9; ++a) ro[a] = buf_rsrc(O.a[E_EX + a], bytes);
    auto rw = [&](int e) { return e < nz ? e : e - nz; };      // e <= nz + 1
    auto load_fields = [&](int r, EmFieldRow<T>& f) {
        const unsigned i = (unsigned)xc + (unsigned)r * (unsigned)nx;
        const unsigned io = (unsigned)xo + (unsigned)r * (unsigned)nx;
        f.ex = ld32(S.a[E_EX], i); f.ey = ld32(S.a[E_EY], i); f.ez = ld32(S.a[E_EZ], i);
        f.bx = ld32(S.a[E_BX], i); f.by = ld32(S.a[E_BY], i); f.bz = ld32(S.a[E_BZ], i);
        f.byo = ld32(S.a[E_BY], io); f.bzo = ld32(S.a[E_BZ], io);
    };
    auto load_fluid = [&](int r, T (&px)[2], T (&py)[2], T (&pz)[2]) {
        const unsigned i = (unsigned)xc + (unsigned)r * (unsigned)nx;
        for (int s = 0; s < 2; ++s) {
            const int b = em_base(s);
            px[s] = s < nsp ? ld32(O.a[b + EV_PX], i) : T(0);
            py[s] = s < nsp ? ld32(O.a[b + EV_PY], i) : T(0);
            pz[s] = s < nsp ? ld32(O.a[b + EV_PZ], i) : T(0);
        }
    };

    EmFieldRow<T> cur, nxt;
    T px[2], py[2], pz[2];
    load_fields(z0, cur);
    load_fluid(z0, px, py, pz);
    T bxm, bym;                                                 // old Bx, By of row e-1
    {
        const unsigned im = (unsigned)xc + (unsigned)zm_per(z0, nz) * (unsigned)nx;
        bxm = ld32(S.a[E_BX], im);
        bym = ld32(S.a[E_BY], im);
    }
    T exp_ = T(0), eyp = T(0), ezp = T(0);                     // new E of row e-1
    T bxp = T(0), byp = T(0), bzp = T(0);                      // old B of row e-1

#pragma unroll 1
    for (int e = z0; e <= z1; ++e) {
        const int rz = rw(e), rzn = rw(e + 1);
        if (e < z1) load_fields(rzn, nxt);
        T pxn[2], pyn[2], pzn[2];
        load_fluid(rzn, pxn, pyn, pzn);

        // ---- Ampere for row e (EMHD_Physics.cu:21-98, currents main_coupled.cu:99-110) ----
        T jx = T(0), jy = T(0), jz = T(0);
        for (int s = 0; s < 2; ++s) {
            if (s >= nsp) break;
            const T q = T(s ? p.q1 : p.q0);
            const T h = T((s ? p.q1 : p.q0) * 0.5);
            const T a = h * (px[s] + lane_from_right(px[s]));
            const T bb = q * py[s];
            const T c = h * (pz[s] + pzn[s]);
            if (s == 0) { jx = a; jy = bb; jz = c; }
            else { jx = jx + a; jy = jy + bb; jz = jz + c; }
        }
        T exn, eyn, ezn;
        {
            const T Bx0 = w ? T(0) : cur.bx, Bxzm = w ? T(0) : bxm;
            T dBy_dz = (cur.by - bym) / dz;
            T curlB_x = -dBy_dz;
            T dBx_dz = (Bx0 - Bxzm) / dz;
            T dBz_dx, dBy_dx;
            if (fwd) {
                dBz_dx = (cur.bzo - cur.bz) / dx;
                dBy_dx = (cur.byo - cur.by) / dx;
            } else {
                dBz_dx = (cur.bz - cur.bzo) / dx;
                dBy_dx = (cur.by - cur.byo) / dx;
            }
            T curlB_y = dBx_dz - dBz_dx;
            T curlB_z = dBy_dx;
            const T ey = w ? T(0) : cur.ey;
            const T ez = w ? T(0) : cur.ez;
            exn = cur.ex + dt * (c2 * curlB_x - jx / eps0);
            eyn = ey + dt * (c2 * curlB_y - jy / eps0);
            ezn = ez + dt * (c2 * curlB_z - jz / eps0);
        }

        // ---- Faraday and the potential for row e-1 ----
        const T eyr = lane_from_right(eyp), ezr = lane_from_right(ezp);
        const T eyl = lane_from_left(eyp), ezl = lane_from_left(ezp);
        if (e > z0) {
            T dEy_dz = (eyn - eyp) / dz;
            T curlE_x = -dEy_dz;
            T dEx_dz = (exn - exp_) / dz;
            T dEz_dx, dEy_dx;
            if (bwd) {
                dEz_dx = (ezp - ezl) / dx;
                dEy_dx = (eyp - eyl) / dx;
            } else {
                dEz_dx = (ezr - ezp) / dx;
                dEy_dx = (eyr - eyp) / dx;
            }
            T curlE_y = dEx_dz - dEz_dx;
            T curlE_z = dEy_dx;
            const T bx = w ? T(0) : bxp;
            const unsigned i = (unsigned)xc + (unsigned)(e - 1) * (unsigned)nx;
            const T ax = ld32(S.a[E_AX], i), ay = ld32(S.a[E_AY], i), az = ld32(S.a[E_AZ], i);
            const unsigned off = store ? i * (unsigned)sizeof(T) : kNoStore;
            buf_st(bx - dt * curlE_x, ro[E_BX - E_EX], off);
            buf_st(byp - dt * curlE_y, ro[E_BY - E_EX], off);
            buf_st(bzp - dt * curlE_z, ro[E_BZ - E_EX], off);
            buf_st(exp_, ro[0], off);
            buf_st(eyp, ro[E_EY - E_EX], off);
            buf_st(ezp, ro[E_EZ - E_EX], off);
            buf_st(ax - dt * exp_, ro[E_AX - E_EX], off);
            buf_st(ay - dt * eyp, ro[E_AY - E_EX], off);
            buf_st(az - dt * ezp, ro[E_AZ - E_EX], off);
        }
        // carry row e
        exp_ = exn; eyp = eyn; ezp = ezn;
        bxp = cur.bx; byp = cur.by; bzp = cur.bz;
        bxm = cur.bx; bym = cur.by;
        if (e < z1) cur = nxt;
        for (int s = 0; s < 2; ++s) { px[s] = pxn[s]; py[s] = pyn[s]; pz[s] = pzn[s]; }
        (void)rz;
    }
}

template <typename T, int TZ>
static void launch_emhd_maxwell_zm_t(const State<T>& S, const State<T>& O, const KParams& p, hipStream_t st) {
    const int nwx = (p.nx + EMZ_OUT - 1) / EMZ_OUT;
    const int nch = (p.nz + TZ - 1) / TZ;
    const int waves = nwx * nch;
    hipLaunchKernelGGL((k_emhd_maxwell_zm<T, TZ>), dim3(round8((waves + 3) / 4)), dim3(256), 0, st, S, O, p, nwx,
                       nch);
}

// variant 0: 64 x 16 tiles; 1 / 2 / 3: z march with 4 / 8 / 16-row chunks
template <typename T>
void launch_emhd_maxwell(const State<T>& S, const State<T>& O, const KParams& p, int variant, hipStream_t st) {
    switch (variant) {
        case 1: launch_emhd_maxwell_zm_t<T, 4>(S, O, p, st); return;
        case 2: launch_emhd_maxwell_zm_t<T, 8>(S, O, p, st); return;
        case 3: launch_emhd_maxwell_zm_t<T, 16>(S, O, p, st); return;
        default: break;
    }
    const int ntx = (p.nx + MX - 1) / MX, ntz = (p.nz + MZ - 1) / MZ;
    dim3 grid(round8(ntx * ntz));
    hipLaunchKernelGGL(k_emhd_maxwell<T>, grid, dim3(MX * MZT), 0, st, S, O, p, ntx, ntz);
}

// --------------------------------------------------------------------------
// Fused x + z sweep, four waves per workgroup.  The workgroup owns 128
// columns of one species and marches a chunk of TZ rows (+3 halo rows each
// side).  Per extended row: all 256 threads stage the 134-position row
// segment, x-sweep it in LDS three variables at a time, and leave the x-swept
// values of the 128 columns in LDS; then waves 0/1 (columns 0-63 / 64-127)
// push n, px, pz and waves 2/3 push Q1, Q2, Q3 into their register FCT
// pipelines along z, with the z sources from row-carried field values.  Only
// the step-start state is read and only the swept fluid is written.
// --------------------------------------------------------------------------
constexpr int XZ2_C = 128;            // columns per workgroup
constexpr int XZ2_P = XZ2_C + 6;      // staged positions (3-cell halo each side)

template <typename T, int TZ>
__global__ __launch_bounds__(256) void k_fct_xz2(State<T> S, State<T> O, KParams p, int ntx, int nch) {
    __shared__ T sU[XZ2_P], sS[XZ2_P], sE[XZ2_P], sNU[XZ2_P], sMU[XZ2_P];
    __shared__ T sV[F_COUNT][XZ2_P];
    __shared__ T sLOR[3][XZ2_P], sLNR[3][XZ2_P], sFL[3][XZ2_P];
    __shared__ T sXS[F_COUNT][XZ2_C];

    int lb = xcd_logical(blockIdx.x, gridDim.x);
    if (lb >= 2 * ntx * nch) return;                 // uniform per block
    const int sp = lb & 1;
    lb >>= 1;
    const int x0 = (lb % ntx) * XZ2_C;
    const int z0 = (lb / ntx) * TZ;
    const int nx = p.nx, nz = p.nz;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int g = wave >> 1;                         // variable group: 0 = n, px, pz; 1 = Q1..Q3
    const int c = (wave & 1) * 64 + lane;            // column within the workgroup
    const int x = x0 + c;
    const bool active = x < nx;
    const int xz = active ? x : nx - 1;              // in-bounds column for loads
    const int xm = xm_src(xz, p);
    const int b = S_NA + sp * F_COUNT;
    const T dt = T(p.dt), dtdx = T(p.dtdx), dtdz = T(p.dtdz);
    const int zend = min(z0 + TZ, nz);
    const int elast = zend + 2;
    const bool wall = !p.periodic_x && (x == 0 || x == nx - 1);

    typedef typename YMModel::template Row<T> Row;
    auto load_row = [&](int r, Row& row) {
        const size_t i = (size_t)xz + (size_t)r * nx, ixm = (size_t)xm + (size_t)r * nx;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (g == 0) {
                row.ez[k] = S.a[S_EZ1 + k][i];
                row.by[k] = S.a[S_BY1 + k][i];
                row.byxm[k] = S.a[S_BY1 + k][ixm];
                row.az[k] = T(0);
            } else {
                row.ez[k] = T(0); row.by[k] = T(0); row.byxm[k] = T(0);
                row.az[k] = S.a[S_AZ1 + k][i];
            }
        }
    };

    int rc = clampi(z0 - 3, 0, nz - 1);
    Row prev, cur;
    load_row(zm_per(rc, nz), prev);
    load_row(rc, cur);
    ZPipe<T, 3> zp;
    zp.Um = T(0);
    zp.MUm = T(0);

    for (int e = z0 - 3; e <= elast; ++e) {
        const size_t row = (size_t)rc * nx;
        // ---- phase A: stage positions q <-> x = x0-3+q ----
        if (tid < XZ2_P) {
            const int q = tid;
            const int xr = map_x(x0 - 3 + q, p);
            const size_t i = row + xr;
            T v[F_COUNT];
#pragma unroll
            for (int k = 0; k < F_COUNT; ++k) v[k] = S.a[b + k][i];
            const T n_safe = kmax(v[F_N], T(0.05));
            sU[q] = v[F_PX] / n_safe;
            sS[q] = lorentz_px(S, sp, xr, rc, p);
#pragma unroll
            for (int k = 0; k < F_COUNT; ++k) sV[k][q] = v[k];
        }
        __syncthreads();
        // ---- phase B: interface coefficients ----
        if (tid < XZ2_P - 1) {
            const int q = tid;
            T vi = T(0.5) * (sU[q] + sU[q + 1]);
            T eps = vi * dtdx;
            T nu, mu;
            fct_nu_mu(eps, nu, mu);
            sE[q] = eps; sNU[q] = nu; sMU[q] = mu;
        }
        __syncthreads();
        // ---- phases C, D, E for variables 0-2, then 3-5 ----
#pragma unroll 1
        for (int v0 = 0; v0 < F_COUNT; v0 += 3) {
            if (tid >= 1 && tid < XZ2_P - 1) {
                const int q = tid;
                const T eL = sE[q - 1], eR = sE[q];
                const T nL = sNU[q - 1], nR = sNU[q];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const int v = v0 + j;
                    const T vm = sV[v][q - 1], vv = sV[v][q], vp = sV[v][q + 1];
                    T fL = T(0.5) * eL * (vm + vv);
                    T fR = T(0.5) * eR * (vv + vp);
                    T dL = nL * (vv - vm);
                    T dR = nR * (vp - vv);
                    T src = (v == F_PX) ? sS[q] : T(0.0);
                    T lor = vv + dt * src - (fR - fL);
                    sLOR[j][q] = lor;
                    sLNR[j][q] = lor + (dR - dL);
                }
            }
            __syncthreads();
            if (tid >= 2 && tid < XZ2_P - 3) {
                const int q = tid;
                const T mu = sMU[q];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    T flx = mu * (sLOR[j][q + 1] - sLOR[j][q]);
                    T l0 = sLNR[j][q - 1], l1 = sLNR[j][q], l2 = sLNR[j][q + 1], l3 = sLNR[j][q + 2];
                    sFL[j][q] = fct_limit(flx, l1 - l0, l2 - l1, l3 - l2);
                }
            }
            __syncthreads();
            if (tid < XZ2_C) {
                const int q = 3 + tid;
#pragma unroll
                for (int j = 0; j < 3; ++j) sXS[v0 + j][tid] = sLNR[j][q] - (sFL[j][q] - sFL[j][q - 1]);
            }
            __syncthreads();
        }
        // ---- z pipeline of this wave's three variables ----
        const int rn = clampi(min(e + 1, elast), 0, nz - 1);
        Row nrow;
        load_row(rn, nrow);
        {
            const int q = 3 + c;
            T F[F_COUNT];
#pragma unroll
            for (int k = 0; k < F_COUNT; ++k) F[k] = sV[k][q];
            T V3[3], S3[3], out[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) V3[j] = sXS[3 * g + j][c];
            const T n_safe = kmax(F[F_N], T(0.05));
            const T u = F[F_PZ] / n_safe;
            const T qq[3] = {F[F_Q1], F[F_Q2], F[F_Q3]};
            if (g == 0) {
                S3[0] = T(0);
                S3[1] = T(0);
                S3[2] = lorentz_pz_v(F[F_N], F[F_PX], qq, cur.ez, prev.ez, cur.by, cur.byxm, prev.by, prev.byxm);
                if (p.warm_T > 0.0) S3[2] -= warm_pz(S, sp, xz, rc, p);
            } else {
                precession_v(F[F_N], F[F_PZ], qq, cur.az, prev.az, T(p.alpha), S3);
            }
            if (zp.push(e, z0, V3, u, S3, dt, dtdz, out) && active) {
                const size_t io = (size_t)x + (size_t)(e - 3) * nx;
                if (g == 0) {
                    if (out[0] < T(0.05)) out[0] = T(0.05);     // kernel_enforce_boundaries
                    if (wall) out[1] = T(0.0);
                }
#pragma unroll
                for (int j = 0; j < 3; ++j) O.a[b + 3 * g + j][io] = out[j];
            }
        }
        YMModel::sel_row(rn != rc, cur, prev);
        cur = nrow;
        rc = rn;
        __syncthreads();                              // LDS is restaged for the next row
    }
}

// --------------------------------------------------------------------------
// Fused x + z sweep in registers (YM).  Lane l of a wavefront owns column
// x = 58*g - 3 + l and marches a chunk of TZ rows (+3 halo rows each side)
// down z exactly as k_fct_z does; for every extended row it first performs
// the x sweep of that row at its column, taking the FCT stencil's neighbours
// (velocity, variables, LORHOT / LNRHOT, limited fluxes, the px source's x-1
// fields) from lanes l-1 / l+1 by DPP shifts, and pushes the x-swept values
// straight into the z pipeline.  The stage never goes to memory and the
// step-start state is read once per row for both sweeps.  The x stencil
// reaches 3 cells, so lanes 0-2 and 61-63 are halo lanes: a wavefront stores
// the 58 columns of lanes 3..60.  A lane whose column lies outside the grid
// loads through map_x and evaluates its px source cell by cell (the source's
// x-1 / x+1 cells follow xm_src / xp_src of the mapped cell, not the lane
// shift), as in k_fct_xr.  Every value is computed by the expressions of
// k_fct_x and k_fct_z in the same order, so the results are those of the
// split sweeps bit for bit.
// --------------------------------------------------------------------------
constexpr int XZR_OUT = 58;

template <typename T, int TZ>
__global__ __launch_bounds__(256) void k_fct_xzr(State<T> S, State<T> O, KParams p, int ngx, int nch) {
    typedef YMModel M;
    constexpr int NV = F_COUNT;
    typedef typename M::template Row<T> Row;
    const int lane = threadIdx.x & 63;
    int wl = xcd_logical(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (wl >= 2 * ngx * nch) return;          // whole wavefronts leave together
    wl = tile_order(wl, 2 * ngx * nch, p.rev);
    const int sp = wl % 2;
    const int g = (wl / 2) % ngx;
    const int ch = wl / (2 * ngx);
    const int nx = p.nx, nz = p.nz;
    const int x = g * XZR_OUT - 3 + lane;      // extended column of this lane
    const int xc = map_x(x, p);                // the real column it reads
    const bool inside = x >= 0 && x < nx;
    const bool store = lane >= 3 && lane < 3 + XZR_OUT && inside;
    const int xm = xm_src(xc, p);
    const bool wall = !p.periodic_x && (xc == 0 || xc == nx - 1);
    const int b = M::base(sp);
    const T dt = T(p.dt), dtdx = T(p.dtdx), dtdz = T(p.dtdz);
    const int z0 = ch * TZ;
    const int zend = min(z0 + TZ, nz);
    const int elast = zend + 2;

    auto load = [&](int r, Row& row, T (&F)[NV], T (&ex)[3]) {
        const unsigned i = (unsigned)xc + (unsigned)r * (unsigned)nx;
        M::load_row(S, i, (unsigned)xm + (unsigned)r * (unsigned)nx, row);
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v] = ld32(S.a[b + v], i);
#pragma unroll
        for (int c = 0; c < 3; ++c) ex[c] = ld32(S.a[S_EX1 + c], i);
    };

    int rc = clampi(z0 - 3, 0, nz - 1);
    Row prev, cur;
    T F[NV], EX[3];
    {
        T Fd[NV], Ed[3];
        load(zm_per(rc, nz), prev, Fd, Ed);
    }
    load(rc, cur, F, EX);
    ZPipe<T, NV> zp;
    zp.Um = T(0);
    zp.MUm = T(0);
    for (int e = z0 - 3; e <= elast; ++e) {
        const int rn = clampi(min(e + 1, elast), 0, nz - 1);
        Row nrow;
        T nF[NV], nEX[3];
        load(rn, nrow, nF, nEX);

        // ---- x sweep of row rc at this lane's column (k_fct_x, YM_FCT.cu:30-91) ----
        // px source (kernel_ym_lorentz, YM_Fluid.cu:32-84) of the cell
        T src;
        {
            const T n_safe = kmax(F[F_N], T(0.05));
            const T vz = F[F_PZ] / n_safe;
            T acc[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T exm = lane_from_left(EX[c]);
                T Exn = T(0.5) * (EX[c] + exm);
                T Byn = T(0.25) * (cur.by[c] + cur.byxm[c] + prev.by[c] + prev.byxm[c]);
                acc[c] = F[F_Q1 + c] * (Exn + vz * Byn);
            }
            src = -(acc[0] + acc[1] + acc[2]);
            if (p.warm_T > 0.0) {
                const T nl = lane_from_left(F[F_N]);
                const T nr = lane_from_right(F[F_N]);
                T dndx = (nr - nl) / (T(2.0) * T(p.dx));
                src -= T(p.warm_T) * dndx;
            }
            if (!inside) src = lorentz_px(S, sp, xc, rc, p);
        }
        T V[NV];
        {
            const T u = F[F_PX] / kmax(F[F_N], T(0.05));          // update_velocities
            const T ur = lane_from_right(u);
            T vi = T(0.5) * (u + ur);
            const T eps = vi * dtdx;
            T nu, mu;
            fct_nu_mu(eps, nu, mu);
            const T epsl = lane_from_left(eps), nul = lane_from_left(nu);
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                const T v0 = F[k];
                const T vm = lane_from_left(v0), vp = lane_from_right(v0);
                T fL = T(0.5) * epsl * (vm + v0);
                T fR = T(0.5) * eps * (v0 + vp);
                T dL = nul * (v0 - vm);
                T dR = nu * (vp - v0);
                T sk = (k == F_PX) ? src : T(0.0);
                T lor = v0 + dt * sk - (fR - fL);
                T lnr = lor + (dR - dL);
                const T lorr = lane_from_right(lor);
                const T lnrl = lane_from_left(lnr);
                const T lnrr = lane_from_right(lnr);
                const T lnrr2 = lane_from_right(lnrr);
                T flx = mu * (lorr - lor);
                T fl = fct_limit(flx, lnr - lnrl, lnrr - lnr, lnrr2 - lnrr);
                const T fll = lane_from_left(fl);
                V[k] = lnr - (fl - fll);
            }
        }

        // ---- z sweep (k_fct_z) of the x-swept row ----
        T uz, Sv[NV], out[NV];
        M::z_src_rows(F, cur, prev, S, sp, xc, rc, p, uz, Sv);
        if (zp.push(e, z0, V, uz, Sv, dt, dtdz, out)) {
            if (out[0] < T(0.05)) out[0] = T(0.05);          // kernel_enforce_boundaries
            if (wall) out[M::IPX] = T(0.0);
            if (store) {
                const unsigned io = (unsigned)xc + (unsigned)(e - 3) * (unsigned)nx;
#pragma unroll
                for (int v = 0; v < NV; ++v) st32_nt(O.a[b + v], io, out[v]);
            }
        }
        M::sel_row(rn != rc, cur, prev);
        cur = nrow;
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v] = nF[v];
#pragma unroll
        for (int c = 0; c < 3; ++c) EX[c] = nEX[c];
        rc = rn;
    }
}

// The same fused march with the variables split into groups on separate
// waves of a workgroup, as k_fct_zg splits the z sweep: a smaller pipeline
// and less x-sweep arithmetic per lane, so more waves fit on a SIMD.  Every
// wave computes the interface coefficients from the cell's n and px and the
// z velocity from n and pz; each reads only the field rows its sources use.
// A group is variables VB .. VB+NGV-1 (FluidVar order) with the rows it
// reads: Ex and By (px source), Ez and By (pz source), Az (precession).
template <int VB_, int NGV_, bool EX_, bool BY_, bool EZ_, bool AZ_>
struct XzGroup {
    static constexpr int VB = VB_, NGV = NGV_;
    static constexpr bool EX = EX_, BY = BY_, EZ = EZ_, AZ = AZ_;
    static constexpr bool has(int k) { return k >= VB && k < VB + NGV; }
};
// two waves: {n, px, pz}, {Q1, Q2, Q3} (a three-wave split {n, px}, {pz},
// {Q} reaches occupancy 3 in FP64 but repeats the shared loads and
// coefficients once more and measured slower)
typedef XzGroup<F_N, 3, true, true, true, false> XzN3;
typedef XzGroup<F_Q1, 3, false, false, false, true> XzQ3;

// WARM: warm_T > 0 (pressure terms in both sources).  EDGE: some lane's
// column lies outside the grid (the first and last column tiles); interior
// wavefronts take the instantiation without the per-cell source fallback, so
// their steady rows are free of branches and the compiler can overlap one
// row's loads and x sweep with the previous row's z pipeline.
template <typename T, int TZ, typename GR, bool WARM, bool EDGE>
__device__ __forceinline__ void xz_march_g(const State<T>& S, const State<T>& O, const KParams& p, int sp, int g,
                                           int ch, int lane) {
    typedef YMModel M;
    constexpr int NV = F_COUNT, NG = GR::NGV, VB = GR::VB;
    typedef typename M::template Row<T> Row;
    const int nx = p.nx, nz = p.nz;
    const int x = g * XZR_OUT - 3 + lane;
    const int xc = map_x(x, p);
    const bool inside = !EDGE || (x >= 0 && x < nx);
    const bool store = lane >= 3 && lane < 3 + XZR_OUT && inside;
    const int xm = xm_src(xc, p);
    const bool wall = !p.periodic_x && (xc == 0 || xc == nx - 1);
    const int b = M::base(sp);
    const T dt = T(p.dt), dtdx = T(p.dtdx), dtdz = T(p.dtdz);
    const int z0 = ch * TZ;
    const int zend = min(z0 + TZ, nz);
    const int elast = zend + 2;
    const unsigned bytes = (unsigned)nx * (unsigned)nz * (unsigned)sizeof(T);
    __amdgpu_buffer_rsrc_t orsc[NG];
#pragma unroll
    for (int j = 0; j < NG; ++j) orsc[j] = buf_rsrc(O.a[b + VB + j], bytes);

    auto load = [&](int r, Row& row, T (&F)[NV], T (&ex)[3]) {
        const unsigned i = (unsigned)xc + (unsigned)r * (unsigned)nx;
        const unsigned ixm = (unsigned)xm + (unsigned)r * (unsigned)nx;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (GR::EZ) row.ez[c] = ld32(S.a[S_EZ1 + c], i);
            if (GR::BY) {
                row.by[c] = ld32(S.a[S_BY1 + c], i);
                // interior wavefronts take By(x-1) from lane l-1 when the
                // row is used (shift_in), not here: a lane shift of a value
                // still being loaded makes the wave wait for that load in
                // the middle of the previous row's arithmetic
                if (EDGE) row.byxm[c] = ld32(S.a[S_BY1 + c], ixm);
            }
            if (GR::AZ) row.az[c] = ld32(S.a[S_AZ1 + c], i);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v] = ld32(S.a[b + v], i);
        if (GR::EX) {
#pragma unroll
            for (int c = 0; c < 3; ++c) ex[c] = ld32(S.a[S_EX1 + c], i);
        }
    };

    // interior wavefronts: By(x-1) of a row from lane l-1 (lane 0's value is
    // unused: it is a halo lane)
    auto shift_in = [&](Row& row) {
        if (!EDGE && GR::BY) {
#pragma unroll
            for (int c = 0; c < 3; ++c) row.byxm[c] = lane_from_left(row.by[c]);
        }
    };

    int rc = clampi(z0 - 3, 0, nz - 1);
    Row prev, cur;
    T F[NV], EX[3];
    {
        T Fd[NV], Ed[3];
        load(zm_per(rc, nz), prev, Fd, Ed);
    }
    load(rc, cur, F, EX);
    shift_in(prev);
    ZPipe<T, NG> zp;
    zp.Um = T(0);
    zp.MUm = T(0);

    // One extended row e (real row rc): load row rn ahead, x-sweep row rc at
    // this lane's column, push it through the z pipeline.  MODE 0: the
    // pipeline's warm-up rows (e < z0 + 3) and the clamped rows at the end of
    // the grid (rn may equal rc); MODE 1: steady rows, where every pipeline
    // stage is active and row rn = rc + 1.
    auto row = [&](int e, auto mode) {
        constexpr int MODE = decltype(mode)::value;
        const int rn = MODE ? rc + 1 : clampi(min(e + 1, elast), 0, nz - 1);
        Row nrow;
        T nF[NV], nEX[3];
        load(rn, nrow, nF, nEX);
        shift_in(cur);

        // ---- x sweep of row rc at this lane's column (YM_FCT.cu:30-91) ----
        T src = T(0.0);
        if (GR::has(F_PX)) {                   // px source (kernel_ym_lorentz, YM_Fluid.cu:32-84)
            const T n_safe = kmax(F[F_N], T(0.05));
            const T vz = F[F_PZ] / n_safe;
            T acc[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T exm = lane_from_left(EX[c]);
                T Exn = T(0.5) * (EX[c] + exm);
                T Byn = T(0.25) * (cur.by[c] + cur.byxm[c] + prev.by[c] + prev.byxm[c]);
                acc[c] = F[F_Q1 + c] * (Exn + vz * Byn);
            }
            src = -(acc[0] + acc[1] + acc[2]);
            if (WARM) {
                const T nl = lane_from_left(F[F_N]);
                const T nr = lane_from_right(F[F_N]);
                T dndx = (nr - nl) / (T(2.0) * T(p.dx));
                src -= T(p.warm_T) * dndx;
            }
            if (EDGE && !inside) src = lorentz_px(S, sp, xc, rc, p);
        }
        T V[NG];
        {
            const T u = F[F_PX] / kmax(F[F_N], T(0.05));          // update_velocities
            const T ur = lane_from_right(u);
            T vi = T(0.5) * (u + ur);
            const T eps = vi * dtdx;
            T nu, mu;
            fct_nu_mu(eps, nu, mu);
#pragma unroll
            for (int j = 0; j < NG; ++j) {
                const int k = VB + j;
                const T v0 = F[k];
                const T vp = lane_from_right(v0);
                // the left interface's fluxes are the left lane's right ones
                // (0.5 eps(x-1) (v(x-1) + v(x)), nu(x-1) (v(x) - v(x-1)):
                // the same operands in the same order)
                T fR = T(0.5) * eps * (v0 + vp);
                T dR = nu * (vp - v0);
                const T fL = lane_from_left(fR);
                const T dL = lane_from_left(dR);
                T sk = (k == F_PX) ? src : T(0.0);
                T lor = v0 + dt * sk - (fR - fL);
                T lnr = lor + (dR - dL);
                const T lorr = lane_from_right(lor);
                const T lnrl = lane_from_left(lnr);
                const T lnrr = lane_from_right(lnr);
                const T lnrr2 = lane_from_right(lnrr);
                T flx = mu * (lorr - lor);
                T fl = fct_limit(flx, lnr - lnrl, lnrr - lnr, lnrr2 - lnrr);
                const T fll = lane_from_left(fl);
                V[j] = lnr - (fl - fll);
            }
        }

        // ---- z sweep of the x-swept row: stale z velocity and the z sources
        // (YMModel::z_src_rows: pz Lorentz force, precession of Q) ----
        T Sv[NG], out[NG];
        const T uz = F[F_PZ] / kmax(F[F_N], T(0.05));
        {
            const T q[3] = {F[F_Q1], F[F_Q2], F[F_Q3]};
            T spz = T(0), sq[3] = {T(0), T(0), T(0)};
            if (GR::has(F_PZ)) {
                spz = lorentz_pz_v(F[F_N], F[F_PX], q, cur.ez, prev.ez, cur.by, cur.byxm, prev.by, prev.byxm);
                if (WARM) spz -= warm_pz(S, sp, xc, rc, p);
            }
            if (GR::has(F_Q1)) precession_v(F[F_N], F[F_PZ], q, cur.az, prev.az, T(p.alpha), sq);
#pragma unroll
            for (int j = 0; j < NG; ++j) {
                const int k = VB + j;
                Sv[j] = (k == F_PZ) ? spz : (k >= F_Q1) ? sq[k >= F_Q1 ? k - F_Q1 : 0] : T(0);
            }
        }
        bool have = true;
        if (MODE) zp.push_steady(V, uz, Sv, dt, dtdz, out);
        else have = zp.push(e, z0, V, uz, Sv, dt, dtdz, out);
        if (have) {
            if constexpr (GR::has(F_N)) {                                        // kernel_enforce_boundaries
                if (out[F_N - VB] < T(0.05)) out[F_N - VB] = T(0.05);
            }
            if constexpr (GR::has(F_PX)) {
                if (wall) out[F_PX - VB] = T(0.0);
            }
            const unsigned off =
                store ? ((unsigned)xc + (unsigned)(e - 3) * (unsigned)nx) * (unsigned)sizeof(T) : kNoStore;
#pragma unroll
            for (int j = 0; j < NG; ++j) buf_st_nt(out[j], orsc[j], off);
        }
        if (MODE) {
            prev = cur;
        } else if (rn != rc) {
            prev = cur;
        }
        cur = nrow;
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v] = nF[v];
        if (GR::EX) {
#pragma unroll
            for (int c = 0; c < 3; ++c) EX[c] = nEX[c];
        }
        rc = rn;
    };

    // rows e = z0-3 .. z0+2 fill the pipeline; steady rows run while e+1 is
    // a real row below elast; the rest (clamped at the grid's end) as warm-up
    int e = z0 - 3;
    for (; e < z0 + 3; ++e) row(e, std::integral_constant<int, 0>());
    const int esteady = min(elast - 1, nz - 2);
    // three rows per iteration: the pipeline's three-row LNRHOT window then
    // returns to the same registers, so it is renamed rather than moved
    for (; e + 2 <= esteady; e += 3) {
        row(e, std::integral_constant<int, 1>());
        row(e + 1, std::integral_constant<int, 1>());
        row(e + 2, std::integral_constant<int, 1>());
    }
    for (; e <= esteady; ++e) row(e, std::integral_constant<int, 1>());
    for (; e <= elast; ++e) row(e, std::integral_constant<int, 0>());
}

template <typename T, int TZ>
__device__ __forceinline__ void xzg_tile(const State<T>& S, const State<T>& O, const KParams& p, int ngx, int wl,
                                         int grp, int lane);

// Workgroup = 2 column groups x 2 variable groups (waves 2i, 2i+1 share the
// columns of tile i, so the second read of the cell's fluid hits the cache).
template <typename T, int TZ, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2))) void k_fct_xzg(State<T> S, State<T> O, KParams p, int ngx,
                                                                                  int nch) {
    constexpr int TPB = BS / 128;             // tiles per workgroup
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int grp = w & 1;
    int wl = __builtin_amdgcn_readfirstlane(xcd_logical(blockIdx.x, gridDim.x) * TPB + (w >> 1));
    if (wl >= 2 * ngx * nch) return;          // whole wavefronts leave together
    wl = tile_order(wl, 2 * ngx * nch, p.rev);
    xzg_tile<T, TZ>(S, O, p, ngx, wl, grp, lane);
}

// Tile wl (uniform) of the fused sweep: species wl % 2, column group
// (wl / 2) % ngx, chunk wl / (2 ngx); variable group grp of the wavefront.
template <typename T, int TZ>
__device__ __forceinline__ void xzg_tile(const State<T>& S, const State<T>& O, const KParams& p, int ngx, int wl,
                                         int grp, int lane) {
    const int sp = wl % 2;
    const int g = (wl / 2) % ngx;
    const int ch = wl / (2 * ngx);
    const bool warm = p.warm_T > 0.0;
    const bool edge = g == 0 || g * XZR_OUT + 60 >= p.nx;   // lanes cover columns 58g-3 .. 58g+60
    if (grp == 0) {
        if (warm) {
            if (edge) xz_march_g<T, TZ, XzN3, true, true>(S, O, p, sp, g, ch, lane);
            else xz_march_g<T, TZ, XzN3, true, false>(S, O, p, sp, g, ch, lane);
        } else {
            if (edge) xz_march_g<T, TZ, XzN3, false, true>(S, O, p, sp, g, ch, lane);
            else xz_march_g<T, TZ, XzN3, false, false>(S, O, p, sp, g, ch, lane);
        }
    } else {
        if (edge) xz_march_g<T, TZ, XzQ3, false, true>(S, O, p, sp, g, ch, lane);   // no warm terms
        else xz_march_g<T, TZ, XzQ3, false, false>(S, O, p, sp, g, ch, lane);
    }
}

// --------------------------------------------------------------------------
// Fused x + z sweep of the EMHD model (n, px, py, pz of one species per
// wavefront), the same register march as xz_march_g: lane l owns column
// 58 g - 3 + l, the x sweep of every row takes its neighbours by DPP and feeds
// the z pipeline.  Both sweeps use the Lorentz force of the step-start state
// (main_coupled.cu:112-136) on px, py, pz, so it is evaluated once per cell
// from the carried field rows.  EDGE wavefronts (some column outside the
// grid) load the x-1 fields through xm_src of the lane's real column, which
// makes the row form equal to em_lorentz at that column for every lane;
// interior wavefronts take Ex, Bz, By at x-1 from lane l-1.
// --------------------------------------------------------------------------
// VB, NG: the variables this wavefront advects (all four, or {n, px} /
// {py, pz} when a species is split over two wavefronts).
template <typename T, int TZ, bool EDGE, int VB = 0, int NG = EV_COUNT>
__device__ __forceinline__ void xz_march_em(const State<T>& S, const State<T>& O, const KParams& p, int sp, int g,
                                            int ch, int lane) {
    constexpr int NV = EV_COUNT;
    typedef EmRow<T> Row;
    const int nx = p.nx, nz = p.nz;
    const int x = g * XZR_OUT - 3 + lane;
    const int xc = map_x(x, p);
    const bool inside = !EDGE || (x >= 0 && x < nx);
    const bool store = lane >= 3 && lane < 3 + XZR_OUT && inside;
    const int xm = xm_src(xc, p);
    const bool wall = !p.periodic_x && (xc == 0 || xc == nx - 1);
    const int b = em_base(sp);
    const T q = T(sp ? p.q1 : p.q0);
    const T dt = T(p.dt), dtdx = T(p.dtdx), dtdz = T(p.dtdz);
    const int z0 = ch * TZ;
    const int zend = min(z0 + TZ, nz);
    const int elast = zend + 2;
    const unsigned bytes = (unsigned)nx * (unsigned)nz * (unsigned)sizeof(T);
    __amdgpu_buffer_rsrc_t orsc[NG];
#pragma unroll
    for (int j = 0; j < NG; ++j) orsc[j] = buf_rsrc(O.a[b + VB + j], bytes);

    auto load = [&](int r, Row& row, T (&F)[NV]) {
        const unsigned i = (unsigned)xc + (unsigned)r * (unsigned)nx;
        const unsigned ixm = (unsigned)xm + (unsigned)r * (unsigned)nx;
        row.ex = ld32(S.a[E_EX], i);
        row.ey = ld32(S.a[E_EY], i);
        row.ez = ld32(S.a[E_EZ], i);
        row.bx = ld32(S.a[E_BX], i);
        row.bz = ld32(S.a[E_BZ], i);
        row.by = ld32(S.a[E_BY], i);
        if (EDGE) {
            row.exm = ld32(S.a[E_EX], ixm);
            row.bzxm = ld32(S.a[E_BZ], ixm);
            row.byxm = ld32(S.a[E_BY], ixm);
        }                                      // interior: shift_in when the row is used
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v] = ld32(S.a[b + v], i);
    };

    // interior wavefronts: the x-1 fields of a row from lane l-1, taken when
    // the row is used (a lane shift of a value still being loaded would make
    // the wave wait for it in the middle of the previous row; lane 0 is a
    // halo lane, its value unused)
    auto shift_in = [&](Row& row) {
        if (!EDGE) {
            row.exm = lane_from_left(row.ex);
            row.bzxm = lane_from_left(row.bz);
            row.byxm = lane_from_left(row.by);
        }
    };

    int rc = clampi(z0 - 3, 0, nz - 1);
    Row prev, cur;
    T F[NV];
    {
        T Fd[NV];
        load(zm_per(rc, nz), prev, Fd);
    }
    load(rc, cur, F);
    shift_in(prev);
    ZPipe<T, NG> zp;
    zp.Um = T(0);
    zp.MUm = T(0);

    auto row = [&](int e, auto mode) {
        constexpr int MODE = decltype(mode)::value;
        const int rn = MODE ? rc + 1 : clampi(min(e + 1, elast), 0, nz - 1);
        Row nrow;
        T nF[NV];
        load(rn, nrow, nF);
        shift_in(cur);

        // Lorentz force of the cell (both sweeps' source on px, py, pz)
        T f[3];
        em_lorentz_v(cur, prev, F[EV_N], F[EV_PX], F[EV_PY], F[EV_PZ], q, f);

        // ---- x sweep of row rc at this lane's column (FCT_Sweeps.cu:184-197) ----
        T V[NG];
        {
            const T u = F[EV_PX] / kmax(F[EV_N], T(0.05));        // update_velocities
            const T ur = lane_from_right(u);
            T vi = T(0.5) * (u + ur);
            const T eps = vi * dtdx;
            T nu, mu;
            fct_nu_mu(eps, nu, mu);
#pragma unroll
            for (int j = 0; j < NG; ++j) {
                const int k = VB + j;
                const T v0 = F[k];
                const T vp = lane_from_right(v0);
                T fR = T(0.5) * eps * (v0 + vp);
                T dR = nu * (vp - v0);
                const T fL = lane_from_left(fR);                    // the left lane's right fluxes
                const T dL = lane_from_left(dR);
                T sk = (k == EV_N) ? T(0.0) : f[k == EV_N ? 0 : k - 1];
                T lor = v0 + dt * sk - (fR - fL);
                T lnr = lor + (dR - dL);
                const T lorr = lane_from_right(lor);
                const T lnrl = lane_from_left(lnr);
                const T lnrr = lane_from_right(lnr);
                const T lnrr2 = lane_from_right(lnrr);
                T flx = mu * (lorr - lor);
                T fl = fct_limit(flx, lnr - lnrl, lnrr - lnr, lnrr2 - lnrr);
                const T fll = lane_from_left(fl);
                V[j] = lnr - (fl - fll);
            }
        }

        // ---- z sweep (FCT_Sweeps.cu:199-232) of the x-swept row ----
        const T uz = F[EV_PZ] / kmax(F[EV_N], T(0.05));
        const T Sall[NV] = {T(0), f[0], f[1], f[2]};
        T Sv[NG], out[NG];
#pragma unroll
        for (int j = 0; j < NG; ++j) Sv[j] = Sall[VB + j];
        bool have = true;
        if (MODE) zp.push_steady(V, uz, Sv, dt, dtdz, out);
        else have = zp.push(e, z0, V, uz, Sv, dt, dtdz, out);
        if (have) {
            if (VB == EV_N && out[0] < T(0.05)) out[0] = T(0.05);          // enforce_boundaries (fluid)
            if constexpr (VB <= EV_PX && EV_PX < VB + NG) {
                if (wall) out[EV_PX - VB] = T(0.0);
            }
            const unsigned off =
                store ? ((unsigned)xc + (unsigned)(e - 3) * (unsigned)nx) * (unsigned)sizeof(T) : kNoStore;
#pragma unroll
            for (int j = 0; j < NG; ++j) buf_st_nt(out[j], orsc[j], off);
        }
        if (MODE) {
            prev = cur;
        } else if (rn != rc) {
            prev = cur;
        }
        cur = nrow;
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v] = nF[v];
        rc = rn;
    };

    int e = z0 - 3;
    for (; e < z0 + 3; ++e) row(e, std::integral_constant<int, 0>());
    const int esteady = min(elast - 1, nz - 2);
    for (; e + 2 <= esteady; e += 3) {
        row(e, std::integral_constant<int, 1>());
        row(e + 1, std::integral_constant<int, 1>());
        row(e + 2, std::integral_constant<int, 1>());
    }
    for (; e <= esteady; ++e) row(e, std::integral_constant<int, 1>());
    for (; e <= elast; ++e) row(e, std::integral_constant<int, 0>());
}

// one wavefront per (species, 58-column tile, z chunk), 4 per workgroup
template <typename T, int TZ>
__global__ __launch_bounds__(256) void k_fct_xze(State<T> S, State<T> O, KParams p, int nsp, int ngx, int nch) {
    const int lane = threadIdx.x & 63;
    int wl = __builtin_amdgcn_readfirstlane(xcd_logical(blockIdx.x, gridDim.x) * 4 + (int)(threadIdx.x >> 6));
    if (wl >= nsp * ngx * nch) return;        // whole wavefronts leave together
    wl = tile_order(wl, nsp * ngx * nch, p.rev);
    const int sp = wl % nsp;
    const int g = (wl / nsp) % ngx;
    const int ch = wl / (nsp * ngx);
    if (g == 0 || g * XZR_OUT + 60 >= p.nx) xz_march_em<T, TZ, true>(S, O, p, sp, g, ch, lane);
    else xz_march_em<T, TZ, false>(S, O, p, sp, g, ch, lane);
}

// the same with each species' variables on two wavefronts ({n, px}, {py, pz}):
// 2 waves per (species, tile, chunk), 4 per workgroup
template <typename T, int TZ>
__global__ __launch_bounds__(256) void k_fct_xze2(State<T> S, State<T> O, KParams p, int nsp, int ngx, int nch) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int wl = __builtin_amdgcn_readfirstlane(xcd_logical(blockIdx.x, gridDim.x) * 2 + (w >> 1));
    if (wl >= nsp * ngx * nch) return;        // both wavefronts of a tile leave together
    wl = tile_order(wl, nsp * ngx * nch, p.rev);
    const int sp = wl % nsp;
    const int g = (wl / nsp) % ngx;
    const int ch = wl / (nsp * ngx);
    const bool edge = g == 0 || g * XZR_OUT + 60 >= p.nx;
    if ((w & 1) == 0) {
        if (edge) xz_march_em<T, TZ, true, EV_N, 2>(S, O, p, sp, g, ch, lane);
        else xz_march_em<T, TZ, false, EV_N, 2>(S, O, p, sp, g, ch, lane);
    } else {
        if (edge) xz_march_em<T, TZ, true, EV_PY, 2>(S, O, p, sp, g, ch, lane);
        else xz_march_em<T, TZ, false, EV_PY, 2>(S, O, p, sp, g, ch, lane);
    }
}

template <typename T, int TZ>
static void launch_fct_xze2_t(const State<T>& S, const State<T>& O, const KParams& p, int nsp, hipStream_t st) {
    const int ngx = (p.nx + XZR_OUT - 1) / XZR_OUT;
    const int nch = (p.nz + TZ - 1) / TZ;
    const int tiles = nsp * ngx * nch;
    hipLaunchKernelGGL((k_fct_xze2<T, TZ>), dim3(round8((tiles + 1) / 2)), dim3(256), 0, st, S, O, p, nsp, ngx,
                       nch);
}

template <typename T, int TZ>
static void launch_fct_xze_t(const State<T>& S, const State<T>& O, const KParams& p, int nsp, hipStream_t st) {
    const int ngx = (p.nx + XZR_OUT - 1) / XZR_OUT;
    const int nch = (p.nz + TZ - 1) / TZ;
    const int waves = nsp * ngx * nch;
    hipLaunchKernelGGL((k_fct_xze<T, TZ>), dim3(round8((waves + 3) / 4)), dim3(256), 0, st, S, O, p, nsp, ngx, nch);
}

// EMHD fused sweep of species 0..nsp-1, variants 0/1/2/3: 16/32/64/8-row
// chunks; 4/5/6: the variables on two wavefronts, 8/16/32-row chunks
template <typename T>
void launch_fct_xz_em(const State<T>& S, const State<T>& O, const KParams& p, int nsp, int variant, hipStream_t st) {
    if (variant == 4) launch_fct_xze2_t<T, 8>(S, O, p, nsp, st);
    else if (variant == 5) launch_fct_xze2_t<T, 16>(S, O, p, nsp, st);
    else if (variant == 6) launch_fct_xze2_t<T, 32>(S, O, p, nsp, st);
    else if (variant == 0) launch_fct_xze_t<T, 16>(S, O, p, nsp, st);
    else if (variant == 2) launch_fct_xze_t<T, 64>(S, O, p, nsp, st);
    else if (variant == 3) launch_fct_xze_t<T, 8>(S, O, p, nsp, st);
    else launch_fct_xze_t<T, 32>(S, O, p, nsp, st);
}
template void launch_fct_xz_em<double>(const State<double>&, const State<double>&, const KParams&, int, int,
                                       hipStream_t);
template void launch_fct_xz_em<float>(const State<float>&, const State<float>&, const KParams&, int, int,
                                      hipStream_t);

template <typename T, int TZ, int BS = 256>
static void launch_fct_xzg_t(const State<T>& S, const State<T>& O, const KParams& p, hipStream_t st) {
    constexpr int TPB = BS / 128;
    const int ngx = (p.nx + XZR_OUT - 1) / XZR_OUT;
    const int nch = (p.nz + TZ - 1) / TZ;
    const int tiles = 2 * ngx * nch;
    hipLaunchKernelGGL((k_fct_xzg<T, TZ, BS>), dim3(round8((tiles + TPB - 1) / TPB)), dim3(BS), 0, st, S, O, p,
                       ngx, nch);
}

template <typename T, int TZ>
static void launch_fct_xzr_t(const State<T>& S, const State<T>& O, const KParams& p, hipStream_t st) {
    const int ngx = (p.nx + XZR_OUT - 1) / XZR_OUT;
    const int nch = (p.nz + TZ - 1) / TZ;
    const int waves = 2 * ngx * nch;
    hipLaunchKernelGGL((k_fct_xzr<T, TZ>), dim3(round8((waves + 3) / 4)), dim3(256), 0, st, S, O, p, ngx, nch);
}

template <typename T, int TZ>
static void launch_fct_xz2_t(const State<T>& S, const State<T>& O, const KParams& p, hipStream_t st) {
    const int ntx = (p.nx + XZ2_C - 1) / XZ2_C;
    const int nch = (p.nz + TZ - 1) / TZ;
    hipLaunchKernelGGL((k_fct_xz2<T, TZ>), dim3(round8(2 * ntx * nch)), dim3(256), 0, st, S, O, p, ntx, nch);
}

// Variant 0: one wave per 64 columns, 32-row chunks; 1/2/3: four waves per
// 128 columns (variable groups split across waves), 16/32/64-row chunks;
// 4/5/6: the x sweep in registers (DPP) inside the z march, 32/64/128-row
// chunks; 7/8/9: the same with the variables on two waves, 16/32/64-row
// chunks (two column tiles per workgroup), 10 / 11 / 12: 32 / 8 / 4-row chunks,
// one tile per workgroup (short chunks: lower latency on small grids); 13 /
// 14: the sweep waves of the Maxwell + next-sweep launch, fed by four waves
// that load the rows into the LDS ring instead of computing Maxwell, chunks
// sized for one / two rounds of resident workgroups (no warm terms).
template <typename T>
void launch_sweep_ring(const State<T>& S, const State<T>& O, const KParams& p, int tz, hipStream_t st);
template <typename T>
void launch_fct_xz(const State<T>& S, const State<T>& O, const KParams& p, int variant, hipStream_t st) {
    switch (variant) {
        case 1: launch_fct_xz2_t<T, 16>(S, O, p, st); return;
        case 2: launch_fct_xz2_t<T, 32>(S, O, p, st); return;
        case 3: launch_fct_xz2_t<T, 64>(S, O, p, st); return;
        case 4: launch_fct_xzr_t<T, 32>(S, O, p, st); return;
        case 5: launch_fct_xzr_t<T, 64>(S, O, p, st); return;
        case 6: launch_fct_xzr_t<T, 128>(S, O, p, st); return;
        case 7: launch_fct_xzg_t<T, 16>(S, O, p, st); return;
        case 8: launch_fct_xzg_t<T, 32>(S, O, p, st); return;
        case 9: launch_fct_xzg_t<T, 64>(S, O, p, st); return;
        case 10: launch_fct_xzg_t<T, 32, 128>(S, O, p, st); return;
        case 11: launch_fct_xzg_t<T, 8, 128>(S, O, p, st); return;
        case 12: launch_fct_xzg_t<T, 4, 128>(S, O, p, st); return;
        case 13: launch_sweep_ring<T>(S, O, p, -1, st); return;
        case 14: launch_sweep_ring<T>(S, O, p, -2, st); return;
        default: break;
    }
    constexpr int TZ = 32;
    const int ntx = (p.nx + XZ_W - 1) / XZ_W;
    const int nch = (p.nz + TZ - 1) / TZ;
    dim3 grid(round8(2 * ntx * nch));
    hipLaunchKernelGGL((k_fct_xz<T, TZ>), grid, dim3(XZ_W), 0, st, S, O, p, ntx, nch);
}

template <typename T, int TZ, int PF>
static int launch_maxwell_zm_t(const State<T>& S, const State<T>& O, const KParams& p, const double* decay,
                               const int* action, hipStream_t st, double* epart = nullptr) {
    const int nwx = (p.nx + MZM_OUT - 1) / MZM_OUT;
    const int nch = (p.nz + TZ - 1) / TZ;
    const int waves = nwx * nch;
    if (epart)
        hipLaunchKernelGGL((k_maxwell_zm<T, TZ, PF, true>), dim3(round8((waves + 3) / 4)), dim3(256), 0, st, S, O, p,
                           nwx, nch, decay, action, epart);
    else
        hipLaunchKernelGGL((k_maxwell_zm<T, TZ, PF>), dim3(round8((waves + 3) / 4)), dim3(256), 0, st, S, O, p, nwx,
                           nch, decay, action, nullptr);
    return waves;
}

// Fused z sweep + Maxwell, variant 1 / 2 / 3: 16 / 32 / 64-row chunks.
template <typename T>
void launch_fct_zmx(const State<T>& S, const Stage<T>& X, const State<T>& O, const KParams& p, const double* decay,
                    const int* action, int variant, hipStream_t st) {
    const int nwx = (p.nx + MZM_OUT - 1) / MZM_OUT;
    const int tz = variant == 1 ? 16 : variant == 3 ? 64 : 32;
    const int nch = (p.nz + tz - 1) / tz;
    dim3 grid(round8(nwx * nch));
    if (tz == 16)
        hipLaunchKernelGGL((k_fct_zmx<T, 16>), grid, dim3(192), 0, st, S, X, O, p, nwx, nch, decay, action);
    else if (tz == 64)
        hipLaunchKernelGGL((k_fct_zmx<T, 64>), grid, dim3(192), 0, st, S, X, O, p, nwx, nch, decay, action);
    else
        hipLaunchKernelGGL((k_fct_zmx<T, 32>), grid, dim3(192), 0, st, S, X, O, p, nwx, nch, decay, action);
}

template <typename T>
void launch_maxwell(const State<T>& S, const State<T>& O, const KParams& p, const double* decay,
                    const int* action, int variant, hipStream_t st) {
    switch (variant) {
        case 4: launch_maxwell_zm_t<T, 16, 1>(S, O, p, decay, action, st); return;
        case 5: launch_maxwell_zm_t<T, 32, 1>(S, O, p, decay, action, st); return;
        case 6: launch_maxwell_zm_t<T, 16, 0>(S, O, p, decay, action, st); return;
        case 7: launch_maxwell_zm_t<T, 8, 1>(S, O, p, decay, action, st); return;
        case 8: launch_maxwell_zm_t<T, 8, 0>(S, O, p, decay, action, st); return;
        case 9: launch_maxwell_zm_t<T, 4, 1>(S, O, p, decay, action, st); return;
        case 10: launch_maxwell_zm_t<T, 2, 1>(S, O, p, decay, action, st); return;
        default: break;
    }
    if (variant == 1) {
        const int ntx = (p.nx + MC_X - 1) / MC_X, ntz = (p.nz + MC_Z - 1) / MC_Z;
        hipLaunchKernelGGL(k_maxwell_cell<T>, dim3(round8(ntx * ntz)), dim3(MC_X * MC_Z), 0, st, S, O, p, ntx, ntz,
                           decay, action);
        return;
    }
    const int tm = (variant == 2) ? 4 : (variant == 3) ? 8 : 16;
    const int ntx = (p.nx + MX - 1) / MX, ntz = (p.nz + tm - 1) / tm;
    dim3 grid(round8(ntx * ntz));
    if (tm == 4)
        hipLaunchKernelGGL((k_maxwell<T, 4>), grid, dim3(MX * MZT), 0, st, S, O, p, ntx, ntz, decay, action);
    else if (tm == 8)
        hipLaunchKernelGGL((k_maxwell<T, 8>), grid, dim3(MX * MZT), 0, st, S, O, p, ntx, ntz, decay, action);
    else
        hipLaunchKernelGGL((k_maxwell<T, 16>), grid, dim3(MX * MZT), 0, st, S, O, p, ntx, ntz, decay, action);
}

// Maxwell as a z march (variants 4..10) that also leaves, in epart, one
// compensated partial sum of the new state's energy density per wavefront
// (the fields after Maxwell and its column action, the fluid it reads), for
// steps whose remaining filters leave the energy unchanged (the charge
// filter).  Returns the number of partial sums, 0 for the other variants
// (nothing launched).
template <typename T>
int launch_maxwell_energy(const State<T>& S, const State<T>& O, const KParams& p, const double* decay,
                          const int* action, int variant, double* epart, hipStream_t st) {
    switch (variant) {
        case 4: return launch_maxwell_zm_t<T, 16, 1>(S, O, p, decay, action, st, epart);
        case 5: return launch_maxwell_zm_t<T, 32, 1>(S, O, p, decay, action, st, epart);
        case 6: return launch_maxwell_zm_t<T, 16, 0>(S, O, p, decay, action, st, epart);
        case 7: return launch_maxwell_zm_t<T, 8, 1>(S, O, p, decay, action, st, epart);
        case 8: return launch_maxwell_zm_t<T, 8, 0>(S, O, p, decay, action, st, epart);
        case 9: return launch_maxwell_zm_t<T, 4, 1>(S, O, p, decay, action, st, epart);
        case 10: return launch_maxwell_zm_t<T, 2, 1>(S, O, p, decay, action, st, epart);
        default: return 0;
    }
}
template int launch_maxwell_energy<double>(const State<double>&, const State<double>&, const KParams&, const double*,
                                           const int*, int, double*, hipStream_t);
template int launch_maxwell_energy<float>(const State<float>&, const State<float>&, const KParams&, const double*,
                                          const int*, int, double*, hipStream_t);

// ==========================================================================
// Maxwell of step n and the fused sweeps of step n + 1 in one launch (YM
// model; only per-column filters -- sponge, xi_cut -- between them).  The
// reference's order is sweeps -> Maxwell -> filters -> the next step's
// sources and sweeps (main_ym.cu:322-455).  Maxwell and the sweeps are local
// in x and z, so one workgroup can do both for a column tile and z chunk:
//   * waves 4 .. 4 + M - 1 compute Maxwell (the arithmetic of k_maxwell_zm,
//     operation for operation) for the rows the tile's sweep reads, on the
//     fields of step n and the fluid after the sweeps of step n, one row per
//     task, the rows dealt round-robin to the M waves; each applies the
//     column action to the colour-2/3 fields, stores the tile's own cells
//     and puts the new field row into an LDS ring of MXZ_SLOTS rows;
//   * waves 0..3 march the fused x + z sweep of xz_march_g (species w / 2,
//     variable group w % 2) on that fluid, with the column action applied to
//     the colour-2/3 charges as they are loaded (the filter stage acts on the
//     charges after Maxwell, whose currents use them unfiltered), and take
//     the new field rows of their sources from the ring.
// The waves are not in lockstep: a Maxwell wave publishes each ring row it
// has written in that row's slot, each sweep wave how many rows it no longer
// reads, and each waits (in LDS, with a bounded spin) only for what it needs,
// so Maxwell runs up to MXZ_SLOTS - 2 rows ahead and a wave's memory stalls
// overlap the others' arithmetic instead of holding the workgroup.  Per cell-update
// the launch reads the fluid and the fields once and writes them once: 47
// arrays, against 36 + 35 for the sweep and Maxwell kernels.
// I: fields of step n, fluid after the sweeps of step n; O: fields of step
// n + 1 (before the charge filter), fluid after the sweeps of step n + 1.
// ==========================================================================
constexpr int MXZ_SLOTS = 8;
// whether the sweep waves take their fluid rows from the ring (filled by the
// Maxwell wave that loads them for its currents) or load them themselves
template <typename T>
constexpr bool kMxzFluidRing = true;

// The ring: slot k % MXZ_SLOTS holds the k-th new field row of the tile
// (state ids S_EX1 .. S_AZ3, one value per lane).  Rows are produced in the
// order the sweep reads them: zm_per(rc0) first (k = 0), then rc0, rc0 + 1,
// ... (k = 1 + row - rc0), rc0 = the chunk's first (clamped) halo row.
// ready[s]: 1 + the ring row whose values slot s holds (0: none yet);
// released[w]: sweep wave w reads no row below it any more.
template <typename T>
struct MxzShared {
    T ring[MXZ_SLOTS][12][64];
    // per species, the sweep's cell velocities p / max(n, 0.05) (x, z) and
    // x-interface coefficients eps, nu, mu of the ring row's fluid
    T vel[MXZ_SLOTS][2][5][64];
    // the ring row's fluid (n, px, pz, Q1..Q3 of both species, before the
    // charge filter; in the sweep-only launch n and pz already swept along
    // x), the sweep waves' input, loaded once by the wave that fills the row
    T fl[kMxzFluidRing<T> ? MXZ_SLOTS : 1][2 * F_COUNT][64];
    int ready[MXZ_SLOTS];
    int released[4];
};
template <typename T>
using MxzRing = T[MXZ_SLOTS][12][64];

// Progress counters in LDS: a relaxed store after a fence that orders this
// wave's LDS accesses before it (writes of the rows it announces, reads of
// the rows it gives up), and a bounded wait followed by the matching fence.
// The fences order LDS only: global loads and stores stay in flight.  The
// bound (2^16 sleeps of 64 cycles, about 2 ms, a thousand times a row's
// work) only guards against a wait that cannot end; with the counters'
// protocol every wait ends.
constexpr int kMxzSpinLimit = 1 << 16;
__device__ __forceinline__ void mxz_publish(int* c, int v) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __hip_atomic_store(c, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int mxz_read(const int* c) {
    return __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// wait until ring rows k and k - 1 are written (their slots' ready counts)
__device__ __forceinline__ void mxz_wait_rows(const int (&ready)[MXZ_SLOTS], int k) {
    for (int i = 0; i < kMxzSpinLimit; ++i) {
        if (mxz_read(&ready[k % MXZ_SLOTS]) >= k + 1 && mxz_read(&ready[(k - 1) % MXZ_SLOTS]) >= k) break;
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
__device__ __forceinline__ void mxz_wait_released(const int (&r)[4], int v) {
    for (int i = 0; i < kMxzSpinLimit; ++i) {
        const int m = min(min(mxz_read(&r[0]), mxz_read(&r[1])), min(mxz_read(&r[2]), mxz_read(&r[3])));
        if (m >= v) break;
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The x sweep of one variable of a row at this lane's column (YM_FCT.cu:30-91):
// v0 its value, sk its source, eps / nu / mu the interface coefficients at
// x + 1/2; every x neighbour from lanes l +- 1, +- 2.
template <typename T>
__device__ __forceinline__ T mxz_xsweep(T v0, T sk, T eps, T nu, T mu, T dt) {
    const T vp = lane_from_right(v0);
    T fR = T(0.5) * eps * (v0 + vp);
    T dR = nu * (vp - v0);
    const T fL = lane_from_left(fR);
    const T dL = lane_from_left(dR);
    T lor = v0 + dt * sk - (fR - fL);
    T lnr = lor + (dR - dL);
    const T lorr = lane_from_right(lor);
    const T lnrl = lane_from_left(lnr);
    const T lnrr = lane_from_right(lnr);
    const T lnrr2 = lane_from_right(lnrr);
    T flx = mu * (lorr - lor);
    T fl = fct_limit(flx, lnr - lnrl, lnrr - lnr, lnrr2 - lnrr);
    const T fll = lane_from_left(fl);
    return lnr - (fl - fll);
}
// Variables whose x sweep the row-loading waves of the sweep-only launch do
// (n and pz of both species: no source, no filter), so that the sweep waves,
// which set the launch's pace, only read the result (2048x1024 FP64: 161-166
// -> 159-161 us; done by the Maxwell waves of the Maxwell + sweep launch
// instead, it makes that launch slower: 223-225 -> 232 us)
__device__ __forceinline__ constexpr bool mxz_xpre(int v) { return v == F_N || v == F_PZ; }

// Waves 4 .. 4 + M - 1: Maxwell as row tasks.  Ring row k is grid row
// r = rf + k (periodic, rf = zm_per(rc0)); Maxwell wave m finishes rows k = m,
// m + M, m + 2M, ..., each from its own loads (rows r - 1 .. r + 1), so the M
// waves work on different rows at once instead of one serial march.  A task
// evaluates Ampere's Ex at rows r and r + 1 (the task of row r + 1 evaluates
// the latter again: the same operations on the same inputs, hence the same
// bits) and, for row r, Jz, Ampere's Ez, the potential and Faraday, with the
// expressions of maxwell_zm_wave in the same order.
template <typename T>
__device__ __forceinline__ void mxz_cell(const MRow<T>& m, T (&ux)[2], T (&uz)[2], T (&nq)[2][3]) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        div2(m.px[s], m.pz[s], kmax(m.n[s], T(1e-10)), ux[s], uz[s]);
#pragma unroll
        for (int c = 0; c < 3; ++c) nq[s][c] = m.n[s] * m.q[s][c];
    }
}
// Ampere's Ex (YM_Physics.cu:70-113) at the row of m; byzm: By of the row
// below, Jx at x + 1/2 from lane l + 1
template <typename T>
__device__ __forceinline__ void mxz_ampere_ex(const MRow<T>& m, const T (&byzm)[3], const T (&ux)[2],
                                              const T (&nq)[2][3], const KParams& p, T (&exn)[3]) {
    const T dt = T(p.dt), aYM = T(p.alpha);
    const CDiv<T> ddz = cdiv_dz<T>(p);
    T Jx[3];
    {
        T vxr[2], nQx[2][3];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            vxr[s] = T(0.5) * (ux[s] + lane_from_right(ux[s]));
#pragma unroll
            for (int c = 0; c < 3; ++c) nQx[s][c] = T(0.5) * (nq[s][c] + lane_from_right(nq[s][c]));
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) Jx[c] = -(nQx[0][c] * vxr[0] + nQx[1][c] * vxr[1]);
    }
    const T cr[3] = {m.az[1] * m.by[2] - m.az[2] * m.by[1], m.az[2] * m.by[0] - m.az[0] * m.by[2],
                     m.az[0] * m.by[1] - m.az[1] * m.by[0]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T dBdz = ddz(m.by[c] - byzm[c]);
        // c^2 (-dBdz - a cr) - Jx / eps0 with c = eps0 = 1 (main_ym.cu:69; the
        // launch is used only then): the product and quotient by 1 are exact
        exn[c] = m.ex[c] + dt * ((-dBdz - aYM * cr[c]) - Jx[c]);
    }
}

template <typename T, int M>
__device__ __forceinline__ void mxz_maxwell(const State<T>& I, const State<T>& O, const KParams& p, int g, int ch,
                                            int tz, int lane, int m, MxzShared<T>& sh, const double* decay,
                                            const int* action) {
    const int nx = p.nx, nz = p.nz;
    const bool per = p.periodic_x != 0;
    const int x = g * XZR_OUT - 3 + lane;
    const int xc = map_x(x, p);
    const int xm = xm_src(xc, p);
    const bool wall = !per && (xc == 0 || xc == nx - 1);     // YM_Fluid.cu:23-26
    const bool last_wall = !per && xc == nx - 1;              // Faraday: E(x+1) := E(x)
    const bool store = lane >= 3 && lane < 3 + XZR_OUT && x >= 0 && x < nx;
    const int z0 = ch * tz;
    const int zend = min(z0 + tz, nz);
    const int elast = zend + 2;
    const int rc0 = clampi(z0 - 3, 0, nz - 1);
    const int rf = zm_per(rc0, nz);
    const T dt = T(p.dt), aYM = T(p.alpha);
    const CDiv<T> ddx = cdiv_dx<T>(p), ddz = cdiv_dz<T>(p);
    const int act = action ? action[xc] : 0;
    const T dec = (act == 1) ? T(decay[xc]) : T(1.0);
    const unsigned bytes = (unsigned)nx * (unsigned)nz * (unsigned)sizeof(T);
    __amdgpu_buffer_rsrc_t rby[3], rex[3], rez[3], raz[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        rby[c] = buf_rsrc(O.a[S_BY1 + c], bytes);
        rex[c] = buf_rsrc(O.a[S_EX1 + c], bytes);
        rez[c] = buf_rsrc(O.a[S_EZ1 + c], bytes);
        raz[c] = buf_rsrc(O.a[S_AZ1 + c], bytes);
    }
    auto rbase = [&](int r) { return (unsigned)((r + nz) % nz) * (unsigned)nx; };
    // rows k = 0 (zm_per(rc0)), 1 (rc0), ..., 1 + clamp(elast) - rc0
    const int nrows = 2 + clampi(elast, 0, nz - 1) - rc0;
    // Finish ring row k (grid row q = rf + k) from its inputs a, the cell
    // quantities of rows q and q + 1 and Ampere's new Ex at both: Jz at
    // (x, q + 1/2), Ampere for Ez, potential, Faraday, the column action,
    // stores, the ring row and the sweep's velocities.
    auto finish = [&](int k, const MRow<T>& a, const T (&uxa)[2], const T (&uza)[2], const T (&nqa)[2][3],
                      const T (&uzb)[2], const T (&nqb)[2][3], const T (&exa)[3], const T (&exb)[3]) {
        const int q = rf + k;
        T ezn[3];
        {
            T vzt[2], nQz[2][3];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                vzt[s] = T(0.5) * (uza[s] + uzb[s]);
#pragma unroll
                for (int c = 0; c < 3; ++c) nQz[s][c] = T(0.5) * (nqa[s][c] + nqb[s][c]);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                T Jz = -(nQz[0][c] * vzt[0] + nQz[1][c] * vzt[1]);
                T dBdx = ddx(a.by[c] - a.byxm[c]);
                T ez = wall ? T(0.0) : a.ez[c];
                ezn[c] = ez + dt * (dBdx - Jz);     // c = eps0 = 1: exact
            }
        }
        T ezr[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            ezr[c] = lane_from_right(ezn[c]);
            if (last_wall) ezr[c] = ezn[c];
        }
        // kernel_ym_potential (YM_Physics.cu:589-599), kernel_ym_faraday (:120-151)
        T Az[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Az[c] = a.az[c];
            if (c > 0 || !p.freeze_az1) Az[c] = Az[c] - dt * ezn[c];
        }
        const T cr[3] = {Az[1] * exa[2] - Az[2] * exa[1], Az[2] * exa[0] - Az[0] * exa[2],
                         Az[0] * exa[1] - Az[1] * exa[0]};
        const int r = (q + nz) % nz;
        // the tile's own cells (each row once: k >= 1)
        const bool own = store && k >= 1 && r >= z0 && r < zend;
        const unsigned off = own ? ((unsigned)r * (unsigned)nx + (unsigned)xc) * (unsigned)sizeof(T) : kNoStore;
        // row k overwrites row k - MXZ_SLOTS: wait until no sweep wave reads it
        if (k >= MXZ_SLOTS) mxz_wait_released(sh.released, k - MXZ_SLOTS + 1);
        const int slot = k % MXZ_SLOTS;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T dEx = ddz(exb[c] - exa[c]);
            T dEz = ddx(ezr[c] - ezn[c]);
            T byn = a.by[c] - dt * (dEx - dEz - aYM * cr[c]);
            T exo = exa[c], ezo = ezn[c], azo = Az[c];
            // sponge / xi_cut of the colour-2/3 fields (YM_Physics.cu:161-210)
            if (c > 0 && act == 1) { byn *= dec; exo *= dec; ezo *= dec; azo *= dec; }
            if (c > 0 && act == 2) { byn = T(0.0); exo = T(0.0); ezo = T(0.0); azo = T(0.0); }
            buf_st(byn, rby[c], off);
            buf_st(exo, rex[c], off);
            buf_st(ezo, rez[c], off);
            if (c > 0 || !p.freeze_az1) buf_st(azo, raz[c], off);
            sh.ring[slot][S_EX1 + c][lane] = exo;
            sh.ring[slot][S_EZ1 + c][lane] = ezo;
            sh.ring[slot][S_BY1 + c][lane] = byn;
            sh.ring[slot][S_AZ1 + c][lane] = azo;
        }
        // the sweep's velocities and x-interface coefficients of row q
        // (update_velocities, YM_Fluid.cu:3-11: p / max(n, 0.05); the
        // currents' quotients above floor n at 1e-10, the same values for
        // n >= 0.05, which every density after a sweep is)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bool low = a.n[s] < T(0.05);
            const T u = low ? a.px[s] / T(0.05) : uxa[s];
            const T uz = low ? a.pz[s] / T(0.05) : uza[s];
            const T ur = lane_from_right(u);
            const T vi = T(0.5) * (u + ur);
            const T eps = vi * T(p.dtdx);
            T nu, mu;
            fct_nu_mu(eps, nu, mu);
            sh.vel[slot][s][0][lane] = u;
            sh.vel[slot][s][1][lane] = uz;
            sh.vel[slot][s][2][lane] = eps;
            sh.vel[slot][s][3][lane] = nu;
            sh.vel[slot][s][4][lane] = mu;
        }
        mxz_publish(&sh.ready[slot], k + 1);
    };

    auto put_fl = [&](int slot, const MRow<T>& r) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            sh.fl[slot][s * F_COUNT + F_N][lane] = r.n[s];
            sh.fl[slot][s * F_COUNT + F_PX][lane] = r.px[s];
            sh.fl[slot][s * F_COUNT + F_PZ][lane] = r.pz[s];
#pragma unroll
            for (int c = 0; c < 3; ++c) sh.fl[slot][s * F_COUNT + F_Q1 + c][lane] = r.q[s][c];
        }
    };

    // Tasks of TR ring rows, dealt round-robin.  FP64: two rows k, k + 1
    // (rows q, q + 1), from loads of rows q - 1 (By) .. q + 2, with Ampere's
    // Ex at q, q + 1, q + 2 (the next task evaluates Ex at its first row
    // q + 2 again: the same operations on the same inputs, the same bits);
    // 2048x1024: 257 -> 247 us.  FP32: one row (two rows' registers spill at
    // four waves per SIMD: 152 -> 161 us).
    constexpr int TR = sizeof(T) == 8 ? 2 : 1;
#pragma unroll 1
    for (int k = TR * m; k < nrows; k += TR * M) {
        const int q = rf + k;                     // unwrapped grid row of ring row k
        const bool two = TR == 2 && k + 1 < nrows;
        MRow<T> a, b, c;                          // rows q, q + 1, q + 2
        T bym[3];                                 // By of row q - 1
        mrow_load(I, I, rbase(q) + (unsigned)xc, rbase(q) + (unsigned)xm, a);
        mrow_load(I, I, rbase(q + 1) + (unsigned)xc, rbase(q + 1) + (unsigned)xm, b);
        if (two) mrow_load(I, I, rbase(q + 2) + (unsigned)xc, rbase(q + 2) + (unsigned)xm, c);
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) bym[cc] = ld32(I.a[S_BY1 + cc], rbase(q - 1) + (unsigned)xc);
        if constexpr (kMxzFluidRing<T>) {
            // the task's fluid rows into their ring slots for the sweep waves,
            // as soon as they are loaded (the slots' rows k - S, k + 1 - S
            // released first), so that they do not stay live in registers
            const int klast = two ? k + 1 : k;
            if (klast >= MXZ_SLOTS) mxz_wait_released(sh.released, klast - MXZ_SLOTS + 1);
            put_fl(k % MXZ_SLOTS, a);
            if (two) put_fl((k + 1) % MXZ_SLOTS, b);
        }
        // cell quantities of the currents (kernel_ym_currents, YM_Physics.cu:9-60)
        T uxa[2], uza[2], nqa[2][3], uxb[2], uzb[2], nqb[2][3];
        mxz_cell(a, uxa, uza, nqa);
        mxz_cell(b, uxb, uzb, nqb);
        T exa[3], exb[3];
        mxz_ampere_ex(a, bym, uxa, nqa, p, exa);
        mxz_ampere_ex(b, a.by, uxb, nqb, p, exb);
        finish(k, a, uxa, uza, nqa, uzb, nqb, exa, exb);
        if constexpr (TR == 2) {
            if (!two) continue;
            T uxc[2], uzc[2], nqc[2][3], exc[3];
            mxz_cell(c, uxc, uzc, nqc);
            mxz_ampere_ex(c, b.by, uxc, nqc, p, exc);
            finish(k + 1, b, uxb, uzb, nqb, uzc, nqc, exb, exc);
        }
    }
}

// Waves 4 .. 4 + M - 1 of the sweep-only launch: ring row k (grid row
// rf + k) from the state's own fields and fluid, no Maxwell; the sweep's
// velocities p / max(n, 0.05) and interface coefficients with it.
template <typename T, int M>
__device__ __forceinline__ void mxz_rows(const State<T>& I, const KParams& p, int g, int ch, int tz, int lane, int m,
                                         MxzShared<T>& sh) {
    const int nx = p.nx, nz = p.nz;
    const int x = g * XZR_OUT - 3 + lane;
    const int xc = map_x(x, p);
    const int xm = xm_src(xc, p);
    const int z0 = ch * tz;
    const int zend = min(z0 + tz, nz);
    const int elast = zend + 2;
    const int rc0 = clampi(z0 - 3, 0, nz - 1);
    const int rf = zm_per(rc0, nz);
    auto rbase = [&](int r) { return (unsigned)((r + nz) % nz) * (unsigned)nx; };
    const int nrows = 2 + clampi(elast, 0, nz - 1) - rc0;
#pragma unroll 1
    for (int k = m; k < nrows; k += M) {
        const int q = rf + k;
        MRow<T> a;
        mrow_load(I, I, rbase(q) + (unsigned)xc, rbase(q) + (unsigned)xm, a);
        if (k >= MXZ_SLOTS) mxz_wait_released(sh.released, k - MXZ_SLOTS + 1);
        const int slot = k % MXZ_SLOTS;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            sh.ring[slot][S_EX1 + c][lane] = a.ex[c];
            sh.ring[slot][S_EZ1 + c][lane] = a.ez[c];
            sh.ring[slot][S_BY1 + c][lane] = a.by[c];
            sh.ring[slot][S_AZ1 + c][lane] = a.az[c];
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            sh.fl[slot][s * F_COUNT + F_N][lane] = a.n[s];
            sh.fl[slot][s * F_COUNT + F_PX][lane] = a.px[s];
            sh.fl[slot][s * F_COUNT + F_PZ][lane] = a.pz[s];
#pragma unroll
            for (int c = 0; c < 3; ++c) sh.fl[slot][s * F_COUNT + F_Q1 + c][lane] = a.q[s][c];
            const T nsafe = kmax(a.n[s], T(0.05));
            const T u = a.px[s] / nsafe;
            const T uz = a.pz[s] / nsafe;
            const T ur = lane_from_right(u);
            const T vi = T(0.5) * (u + ur);
            const T eps = vi * T(p.dtdx);
            T nu, mu;
            fct_nu_mu(eps, nu, mu);
            sh.vel[slot][s][0][lane] = u;
            sh.vel[slot][s][1][lane] = uz;
            sh.vel[slot][s][2][lane] = eps;
            sh.vel[slot][s][3][lane] = nu;
            sh.vel[slot][s][4][lane] = mu;
            sh.fl[slot][s * F_COUNT + F_N][lane] = mxz_xsweep(a.n[s], T(0.0), eps, nu, mu, T(p.dt));
            sh.fl[slot][s * F_COUNT + F_PZ][lane] = mxz_xsweep(a.pz[s], T(0.0), eps, nu, mu, T(p.dt));
        }
        mxz_publish(&sh.ready[slot], k + 1);
    }
}

// Waves 0..3: xz_march_g on the fluid of I with the new field rows from the
// ring.  Lane l reads the fields of its column at ring lane lx and those of
// the source stencil's x-1 column at lxm: l and l-1, except for lanes beyond
// a wall, which take their clamped column's lane (so their sources are
// lorentz_px / lorentz_pz / precession at that column, as the reference's
// clamped halo loads give).
template <typename T, typename GR, bool EDGE, bool XP>
__device__ __forceinline__ void mxz_sweep(const State<T>& I, const State<T>& O, const KParams& p, int sp, int g,
                                          int ch, int tz, int lane, MxzShared<T>& sh, int w, const double* decay,
                                          const int* action) {
    const MxzRing<T>& ring = sh.ring;
    typedef YMModel M;
    constexpr int NV = F_COUNT, NG = GR::NGV, VB = GR::VB;
    const int nx = p.nx, nz = p.nz;
    const int x = g * XZR_OUT - 3 + lane;
    const int xc = map_x(x, p);
    const bool inside = !EDGE || (x >= 0 && x < nx);
    const bool store = lane >= 3 && lane < 3 + XZR_OUT && inside;
    int lx = lane, lxm = lane - 1;
    if (EDGE && !p.periodic_x) {
        lx = lane + (xc - x);
        lxm = xc > 0 ? lx - 1 : lx;
    }
    lxm = lxm < 0 ? 0 : lxm;                    // lane 0's x-1 values are not used
    const bool wall = !p.periodic_x && (xc == 0 || xc == nx - 1);
    const int b = M::base(sp);
    const T dt = T(p.dt), dtdz = T(p.dtdz);
    const int z0 = ch * tz;
    const int zend = min(z0 + tz, nz);
    const int elast = zend + 2;
    const int rc0 = clampi(z0 - 3, 0, nz - 1);
    // sponge / xi_cut of the charges Q2, Q3 (the filter stage of step n)
    const bool cut = action != nullptr;
    const int act = cut ? action[xc] : 0;
    const T dec = (act == 1) ? T(decay[xc]) : T(1.0);
    const unsigned bytes = (unsigned)nx * (unsigned)nz * (unsigned)sizeof(T);
    __amdgpu_buffer_rsrc_t orsc[NG];
#pragma unroll
    for (int j = 0; j < NG; ++j) orsc[j] = buf_rsrc(O.a[b + VB + j], bytes);

    // FP32: the fluid rows come from the ring (the Maxwell wave's loads of
    // the same rows); FP64: loaded two rows ahead, F (row rc(e)), nF
    // (rc(e + 1)), nnF (rc(e + 2)); rc(e) = clamp(min(e, elast), 0, nz - 1)
    constexpr bool FL = kMxzFluidRing<T>;
    auto load = [&](int r, T (&F)[NV]) {
        const unsigned i = (unsigned)xc + (unsigned)r * (unsigned)nx;
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v] = ld32(I.a[b + v], i);
    };
    auto rowof = [&](int e) { return clampi(min(e, elast), 0, nz - 1); };
    int rc = rc0;
    T F[NV], nF[NV];
    if constexpr (!FL) {
        load(rc, F);
        load(rowof(z0 - 2), nF);
    }
    ZPipe<T, NG> zp;
    zp.Um = T(0);
    zp.MUm = T(0);

    auto row = [&](int e, auto mode) {
        constexpr int MODE = decltype(mode)::value;
        const int rn = MODE ? rc + 1 : rowof(e + 1);
        T nnF[NV];
        if constexpr (!FL) load(rowof(e + 2), nnF);
        const int kc = 1 + rc - rc0;            // ring index of row rc; row rc's predecessor is kc - 1
        mxz_wait_rows(sh.ready, kc);
        if constexpr (FL) {
#pragma unroll
            for (int v = 0; v < NV; ++v) F[v] = sh.fl[kc % MXZ_SLOTS][sp * F_COUNT + v][lane];
        }
        if (cut) {
            // k_cell_filters on the charges of the row in use: multiply by
            // the column's decay, or zero
            F[F_Q2] = (act == 2) ? T(0.0) : F[F_Q2] * dec;
            F[F_Q3] = (act == 2) ? T(0.0) : F[F_Q3] * dec;
        }
        const T(&cs)[12][64] = ring[kc % MXZ_SLOTS];
        const T(&ps)[12][64] = ring[(kc + MXZ_SLOTS - 1) % MXZ_SLOTS];
        const T(&vs)[5][64] = sh.vel[kc % MXZ_SLOTS][sp];
        const T u_x = vs[0][lane], u_z = vs[1][lane];

        // ---- x sweep of row rc at this lane's column (YM_FCT.cu:30-91) ----
        T src = T(0.0);
        if (GR::has(F_PX)) {                   // px source (kernel_ym_lorentz, YM_Fluid.cu:32-84)
            const T vz = u_z;
            T acc[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                T Exn = T(0.5) * (cs[S_EX1 + c][lx] + cs[S_EX1 + c][lxm]);
                T Byn = T(0.25) * (cs[S_BY1 + c][lx] + cs[S_BY1 + c][lxm] + ps[S_BY1 + c][lx] + ps[S_BY1 + c][lxm]);
                acc[c] = F[F_Q1 + c] * (Exn + vz * Byn);
            }
            src = -(acc[0] + acc[1] + acc[2]);
        }
        T V[NG];
        {
            // velocity and interface coefficients from the Maxwell wave
            const T eps = vs[2][lane], nu = vs[3][lane], mu = vs[4][lane];
#pragma unroll
            for (int j = 0; j < NG; ++j) {
                const int k = VB + j;
                // n, pz: swept by the wave that filled the row (the ring holds the result)
                if (XP && mxz_xpre(k)) V[j] = F[k];
                else V[j] = mxz_xsweep(F[k], (k == F_PX) ? src : T(0.0), eps, nu, mu, dt);
            }
        }

        // ---- z sweep of the x-swept row: stale z velocity and the z sources ----
        T Sv[NG], out[NG];
        const T uz = u_z;
        {
            const T q[3] = {F[F_Q1], F[F_Q2], F[F_Q3]};
            T spz = T(0), sq[3] = {T(0), T(0), T(0)};
            if (GR::has(F_PZ)) {
                T ez[3], ezm[3], by[3], byxm[3], byzm[3], byxmzm[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    ez[c] = cs[S_EZ1 + c][lx];
                    ezm[c] = ps[S_EZ1 + c][lx];
                    by[c] = cs[S_BY1 + c][lx];
                    byxm[c] = cs[S_BY1 + c][lxm];
                    byzm[c] = ps[S_BY1 + c][lx];
                    byxmzm[c] = ps[S_BY1 + c][lxm];
                }
                spz = lorentz_pz_u(u_x, q, ez, ezm, by, byxm, byzm, byxmzm);
            }
            if (GR::has(F_Q1)) {
                T az[3], azm[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    az[c] = cs[S_AZ1 + c][lx];
                    azm[c] = ps[S_AZ1 + c][lx];
                }
                precession_u(u_z, q, az, azm, T(p.alpha), sq);
            }
#pragma unroll
            for (int j = 0; j < NG; ++j) {
                const int k = VB + j;
                Sv[j] = (k == F_PZ) ? spz : (k >= F_Q1) ? sq[k >= F_Q1 ? k - F_Q1 : 0] : T(0);
            }
        }
        bool have = true;
        if (MODE) zp.push_steady(V, uz, Sv, dt, dtdz, out);
        else have = zp.push(e, z0, V, uz, Sv, dt, dtdz, out);
        if (have) {
            if constexpr (GR::has(F_N)) {                                        // kernel_enforce_boundaries
                if (out[F_N - VB] < T(0.05)) out[F_N - VB] = T(0.05);
            }
            if constexpr (GR::has(F_PX)) {
                if (wall) out[F_PX - VB] = T(0.0);
            }
            const unsigned off =
                store ? ((unsigned)xc + (unsigned)(e - 3) * (unsigned)nx) * (unsigned)sizeof(T) : kNoStore;
#pragma unroll
            for (int j = 0; j < NG; ++j) buf_st_nt(out[j], orsc[j], off);
        }
        if constexpr (!FL) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                F[v] = nF[v];
                nF[v] = nnF[v];
            }
        }
        // the next row reads ring rows kn - 1 and kn: the ones below are free
        mxz_publish(&sh.released[w], kc + (rn != rc) - 1);
        rc = rn;
    };

    int e = z0 - 3;
    for (; e < z0 + 3; ++e) row(e, std::integral_constant<int, 0>());
    const int esteady = min(elast - 1, nz - 2);
    for (; e + 2 <= esteady; e += 3) {
        row(e, std::integral_constant<int, 1>());
        row(e + 1, std::integral_constant<int, 1>());
        row(e + 2, std::integral_constant<int, 1>());
    }
    for (; e <= esteady; ++e) row(e, std::integral_constant<int, 1>());
    for (; e <= elast; ++e) row(e, std::integral_constant<int, 0>());
}

// Workgroup: waves 0..3 sweep (species w / 2, group w % 2), waves 4.. Maxwell;
// tile = 58 columns x a chunk of tz rows.  Occupancy: two waves per SIMD in
// FP64, four in FP32 (three with M = 2: two workgroups' LDS per CU); 4 + M
// waves per workgroup (M = 4: the eight wave slots of a CU in FP64).
template <typename T, int M, bool MX = true>
__global__ __launch_bounds__(64 * (4 + M)) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? (M == 2 ? 3 : 4) : 2))) void k_mxz(
    State<T> I, State<T> O, KParams p, int ngx, int nch, int tz, const double* decay, const int* action) {
    __shared__ MxzShared<T> sh;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int lb = xcd_logical(blockIdx.x, gridDim.x);
    if (lb >= ngx * nch) return;                // whole workgroups leave together
    lb = __builtin_amdgcn_readfirstlane(tile_order(lb, ngx * nch, p.rev));
    const int g = lb % ngx, ch = lb / ngx;
    if (threadIdx.x < MXZ_SLOTS) sh.ready[threadIdx.x] = 0;
    if (threadIdx.x < 4) sh.released[threadIdx.x] = 0;
    __syncthreads();                            // the only barrier: the counters are set
    if (w >= 4) {
        if constexpr (MX) mxz_maxwell<T, M>(I, O, p, g, ch, tz, lane, w - 4, sh, decay, action);
        else mxz_rows<T, M>(I, p, g, ch, tz, lane, w - 4, sh);
        return;
    }
    const int sp = w >> 1;
    const bool edge = g == 0 || g * XZR_OUT + 60 >= p.nx;   // lanes cover columns 58g-3 .. 58g+60
    if (w & 1) {
        if (edge) mxz_sweep<T, XzQ3, true, !MX>(I, O, p, sp, g, ch, tz, lane, sh, w, decay, action);
        else mxz_sweep<T, XzQ3, false, !MX>(I, O, p, sp, g, ch, tz, lane, sh, w, decay, action);
    } else {
        if (edge) mxz_sweep<T, XzN3, true, !MX>(I, O, p, sp, g, ch, tz, lane, sh, w, decay, action);
        else mxz_sweep<T, XzN3, false, !MX>(I, O, p, sp, g, ch, tz, lane, sh, w, decay, action);
    }
}

// Workgroups of k_mxz<T, M> the current device holds at once (its compute
// units times the workgroups per unit that the kernel's registers and LDS
// allow), queried once per device.
template <typename T, int M, bool MX = true>
static int mxz_resident() {
    static std::atomic<int> cache[16];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) dev = 0;
    int r = cache[dev].load(std::memory_order_relaxed);
    if (r == 0) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mxz<T, M, MX>, 64 * (4 + M), 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
        r = per_cu * cus;
        cache[dev].store(r, std::memory_order_relaxed);
    }
    return r;
}

// tz > 0: chunks of tz rows; tz = -R: the chunk height that makes the tiles
// fill the device's resident workgroups R times (at least 16 rows), so that
// the launch runs in R rounds with long chunks instead of many rounds of
// short ones and a partial last round.
template <typename T, int M, bool MX = true>
static void launch_mxz_t(const State<T>& I, const State<T>& O, const KParams& p, int tz, const double* decay,
                         const int* action, hipStream_t st) {
    const int ngx = (p.nx + XZR_OUT - 1) / XZR_OUT;
    if (tz <= 0) {
        const int nch = std::max(1, (-tz * mxz_resident<T, M, MX>()) / ngx);
        tz = std::max(16, (p.nz + nch - 1) / nch);
    }
    const int nch = (p.nz + tz - 1) / tz;
    hipLaunchKernelGGL((k_mxz<T, M, MX>), dim3(round8(ngx * nch)), dim3(64 * (4 + M)), 0, st, I, O, p, ngx, nch,
                       tz, decay, action);
}

// variants 0/1/2: 16/32/64-row chunks with two Maxwell waves, 3/4/5 with
// four; 6/7: four Maxwell waves, chunks sized for one / two rounds of
// resident workgroups; 8: two Maxwell waves, one round
template <typename T>
void launch_maxwell_sweep(const State<T>& I, const State<T>& O, const KParams& p, const double* decay,
                          const int* action, int variant, hipStream_t st) {
    switch (variant) {
        case 0: launch_mxz_t<T, 2>(I, O, p, 16, decay, action, st); return;
        case 2: launch_mxz_t<T, 2>(I, O, p, 64, decay, action, st); return;
        case 3: launch_mxz_t<T, 4>(I, O, p, 16, decay, action, st); return;
        case 4: launch_mxz_t<T, 4>(I, O, p, 32, decay, action, st); return;
        case 5: launch_mxz_t<T, 4>(I, O, p, 64, decay, action, st); return;
        case 6: launch_mxz_t<T, 4>(I, O, p, -1, decay, action, st); return;
        case 7: launch_mxz_t<T, 4>(I, O, p, -2, decay, action, st); return;
        case 8: launch_mxz_t<T, 2>(I, O, p, -1, decay, action, st); return;
        default: launch_mxz_t<T, 2>(I, O, p, 32, decay, action, st); return;
    }
}

// The fused sweeps alone (xz variants 13 / 14): state S -> fluid of O.
template <typename T>
void launch_sweep_ring(const State<T>& S, const State<T>& O, const KParams& p, int tz, hipStream_t st) {
    launch_mxz_t<T, 4, false>(S, O, p, tz, nullptr, nullptr, st);
}

template void launch_maxwell_sweep<double>(const State<double>&, const State<double>&, const KParams&, const double*,
                                           const int*, int, hipStream_t);
template void launch_maxwell_sweep<float>(const State<float>&, const State<float>&, const KParams&, const double*,
                                          const int*, int, hipStream_t);

#define KHS_SWEEP_INST(T, M)                                                                             \
    template void launch_fct_x<T, M>(const State<T>&, const Stage<T>&, const KParams&, int, int, int, hipStream_t); \
    template void launch_fct_z<T, M>(const State<T>&, const Stage<T>&, const State<T>&, const KParams&, int, int, int, hipStream_t);
KHS_SWEEP_INST(double, YMModel)
KHS_SWEEP_INST(float, YMModel)
KHS_SWEEP_INST(double, EMModel)
KHS_SWEEP_INST(float, EMModel)
#undef KHS_SWEEP_INST
// --------------------------------------------------------------------------
// Packed runs (see PackedRun in khs_launch.h): the fused sweep and the
// Maxwell march of many runs in one launch each.  Tile t of a launch belongs
// to the last run r with begin[r] <= t; the workgroup's wavefronts (uniform
// r) take that run's state pointers and parameters into scalar registers and
// run the single-run code on its tile.
// --------------------------------------------------------------------------
__device__ __forceinline__ int packed_run(const int* __restrict__ begin, int n, int t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

template <typename T, int TZ, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 2))) void k_fct_xzg_packed(
    const PackedRun<T>* __restrict__ runs, const int* __restrict__ begin, int n, int total) {
    constexpr int TPB = BS / 128;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int t = __builtin_amdgcn_readfirstlane(xcd_logical(blockIdx.x, gridDim.x) * TPB + (w >> 1));
    if (t >= total) return;                   // whole wavefronts leave together
    const int r = packed_run(begin, n, t);
    // the run's descriptor is read in place (uniform loads of read-only
    // kernel-argument memory): a local copy of the pointer tables would be
    // indexed by the species at run time and live in scratch memory
    const PackedRun<T>& d = runs[r];
    xzg_tile<T, TZ>(d.S, d.O, d.p, d.ngx, t - begin[r], w & 1, lane);
}

template <typename T, int TZ, int PF>
__global__ __launch_bounds__(256) void k_maxwell_zm_packed(const PackedRun<T>* __restrict__ runs,
                                                          const int* __restrict__ begin, int n, int total) {
    const int lane = threadIdx.x & 63;
    const int t = __builtin_amdgcn_readfirstlane(xcd_logical(blockIdx.x, gridDim.x) * 4 + (int)(threadIdx.x >> 6));
    if (t >= total) return;
    const int r = packed_run(begin, n, t);
    const PackedRun<T>& d = runs[r];
    maxwell_zm_wave<T, TZ, PF>(d.S, d.O, d.p, d.nwx, t - begin[r], lane, d.decay, d.action);
}

template <typename T, int XZ, int MX, int BS = 256>
static void launch_packed_step_t(const PackedRun<T>* runs, const int* xz_begin, const int* mx_begin,
                                 const int* f_begin, int n, int xz_total, int mx_total, int f_total, hipStream_t st) {
    constexpr int TPB = BS / 128;
    hipLaunchKernelGGL((k_fct_xzg_packed<T, XZ, BS>), dim3(round8((xz_total + TPB - 1) / TPB)), dim3(BS), 0, st,
                       runs, xz_begin, n, xz_total);
    hipLaunchKernelGGL((k_maxwell_zm_packed<T, MX, 1>), dim3(round8((mx_total + 3) / 4)), dim3(256), 0, st, runs,
                       mx_begin, n, mx_total);
    launch_filters_packed<T>(runs, f_begin, n, f_total, st);
}

template <typename T>
void launch_packed_step(const PackedRun<T>* runs, const int* xz_begin, const int* mx_begin, const int* f_begin,
                        int n, int xz_total, int mx_total, int f_total, int shape, hipStream_t st) {
    switch (shape) {
        case 1: launch_packed_step_t<T, 32, 4>(runs, xz_begin, mx_begin, f_begin, n, xz_total, mx_total, f_total, st); return;
        case 2: launch_packed_step_t<T, 32, 8>(runs, xz_begin, mx_begin, f_begin, n, xz_total, mx_total, f_total, st); return;
        case 3: launch_packed_step_t<T, 16, 4>(runs, xz_begin, mx_begin, f_begin, n, xz_total, mx_total, f_total, st); return;
        case 4: launch_packed_step_t<T, 32, 2>(runs, xz_begin, mx_begin, f_begin, n, xz_total, mx_total, f_total, st); return;
        case 5: launch_packed_step_t<T, 16, 2>(runs, xz_begin, mx_begin, f_begin, n, xz_total, mx_total, f_total, st); return;
        case 6: launch_packed_step_t<T, 32, 4, 128>(runs, xz_begin, mx_begin, f_begin, n, xz_total, mx_total, f_total, st); return;
        case 7: launch_packed_step_t<T, 8, 4, 128>(runs, xz_begin, mx_begin, f_begin, n, xz_total, mx_total, f_total, st); return;
        default: launch_packed_step_t<T, 16, 8>(runs, xz_begin, mx_begin, f_begin, n, xz_total, mx_total, f_total, st); return;
    }
}
template void launch_packed_step<float>(const PackedRun<float>*, const int*, const int*, const int*, int, int, int,
                                        int, int, hipStream_t);
template void launch_packed_step<double>(const PackedRun<double>*, const int*, const int*, const int*, int, int,
                                         int, int, int, hipStream_t);

template void launch_emhd_maxwell<double>(const State<double>&, const State<double>&, const KParams&, int,
                                          hipStream_t);
template void launch_emhd_maxwell<float>(const State<float>&, const State<float>&, const KParams&, int, hipStream_t);
template void launch_fct_xz<double>(const State<double>&, const State<double>&, const KParams&, int, hipStream_t);
template void launch_fct_xz<float>(const State<float>&, const State<float>&, const KParams&, int, hipStream_t);
template void launch_fct_zmx<double>(const State<double>&, const Stage<double>&, const State<double>&, const KParams&,
                                     const double*, const int*, int, hipStream_t);
template void launch_fct_zmx<float>(const State<float>&, const Stage<float>&, const State<float>&, const KParams&,
                                    const double*, const int*, int, hipStream_t);
template void launch_maxwell<double>(const State<double>&, const State<double>&, const KParams&, const double*, const int*, int, hipStream_t);
template void launch_maxwell<float>(const State<float>&, const State<float>&, const KParams&, const double*, const int*, int, hipStream_t);

}  // namespace khs
