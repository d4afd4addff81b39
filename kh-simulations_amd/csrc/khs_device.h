// khs_device.h — device-side types and inline physics shared by the gfx950 kernels.
//
// All kernels are templated on the storage/arithmetic type T (double for the
// FP64 drop-in, float for the precision form).  The FP64 expressions keep the
// reference's evaluation order (citations inline) and the library is built with
// -ffp-contract=off, so FP64 stencil results are reproducible bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "khs_layout.h"

namespace khs {

template <typename T>
struct State {
    T* a[S_COUNT];
};

template <typename T>
struct Stage {            // x-swept fluid of both species, 12 arrays
    T* a[2 * F_COUNT];
};

// Scalar parameters of one run, passed by value to every kernel.
struct KParams {
    int nx, nz;
    int periodic_x;
    int freeze_az1;
    double dt, dx, dz;
    double dtdx, dtdz;        // dt/dx, dt/dz (the reference evaluates dt_step/dx per call; same value)
    double c2, eps0, alpha, warm_T;
    double rdx, rdz, reps0;   // 1/dx, 1/dz, 1/eps0 rounded to nearest (CDiv)
    double q0, q1;            // EMHD species charges (emhdgpu2d fluid: -1)
    int nsp;                  // EMHD: 1 or 2 fluids (YM: always 2)
    int rev;                  // walk the launch's tiles in reverse logical order
};

// Division by a launch-wide constant b > 0 without a division sequence
// (FP64): y = RN(1/b) is computed on the host, q = RN(a y), then two
// corrections q = RN(q + r y) with the remainder r = a - b q, exact in one
// fma.  After the first correction q is within one ulp of a / b; with y
// within half an ulp of 1/b the second gives RN(a / b) (Markstein's theorem,
// IBM J. Res. Dev. 34(1), 1990): the bits of IEEE division, zeros of either
// sign included (the remainder is formed as b q - a and subtracted), while
// the remainder stays normal (|a| above ~2^-960).  5 FP64 operations instead
// of 11 with a quarter-rate reciprocal.  The FP32 form divides: there the
// remainder of |a| below ~2^-100 is subnormal and the correction can miss by
// one ulp.
template <typename T>
struct CDiv {
    T b, y;
    __device__ __forceinline__ T operator()(T a) const {
        if constexpr (sizeof(T) == 4) {
            return a / b;
        } else {
            T q = a * y;
            T rn = __builtin_fma(b, q, -a);
            q = __builtin_fma(-rn, y, q);
            rn = __builtin_fma(b, q, -a);
            return __builtin_fma(-rn, y, q);
        }
    }
};
template <typename T> __device__ __forceinline__ CDiv<T> cdiv_dx(const KParams& p);
template <> __device__ __forceinline__ CDiv<double> cdiv_dx<double>(const KParams& p) { return {p.dx, p.rdx}; }
template <> __device__ __forceinline__ CDiv<float> cdiv_dx<float>(const KParams& p) { return {float(p.dx), 0.0f}; }
template <typename T> __device__ __forceinline__ CDiv<T> cdiv_dz(const KParams& p);
template <> __device__ __forceinline__ CDiv<double> cdiv_dz<double>(const KParams& p) { return {p.dz, p.rdz}; }
template <> __device__ __forceinline__ CDiv<float> cdiv_dz<float>(const KParams& p) { return {float(p.dz), 0.0f}; }
template <typename T> __device__ __forceinline__ CDiv<T> cdiv_eps0(const KParams& p);
template <> __device__ __forceinline__ CDiv<double> cdiv_eps0<double>(const KParams& p) { return {p.eps0, p.reps0}; }
template <> __device__ __forceinline__ CDiv<float> cdiv_eps0<float>(const KParams& p) { return {float(p.eps0), 0.0f}; }

// a1 / b and a2 / b with one reciprocal (FP64): the operations of the
// compiler's IEEE double division on gfx950 (reciprocal estimate v_rcp_f64,
// two Newton steps, the quotient a y and one fma correction by the exact
// remainder) with the reciprocal shared, and without that sequence's range
// scaling and special-case fix-up, which are identities for finite operands
// whose magnitudes and quotients lie well inside the normal range (momenta
// over floored densities here).  The remainder is formed as b q - a and
// subtracted, so a zero numerator keeps its sign.  The FP32 form divides.
__device__ __forceinline__ void div2(double a1, double a2, double b, double& q1, double& q2) {
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    double q = a1 * y;
    q1 = __builtin_fma(-__builtin_fma(b, q, -a1), y, q);
    q = a2 * y;
    q2 = __builtin_fma(-__builtin_fma(b, q, -a2), y, q);
}
__device__ __forceinline__ void div2(float a1, float a2, float b, float& q1, float& q2) {
    q1 = a1 / b;
    q2 = a2 / b;
}

// Compensated FP64 sums: Neumaier accumulation and TwoSum merges.
struct CSum {
    double s, c;
};
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void cadd(CSum& acc, double v) {   // Neumaier
    double s, e;
    two_sum(acc.s, v, s, e);
    acc.s = s;
    acc.c += e;
}
__device__ __forceinline__ CSum cmerge(CSum a, CSum b) {
    double s, e;
    two_sum(a.s, b.s, s, e);
    return CSum{s, a.c + b.c + e};
}

// --------------------------------------------------------------------------
// index helpers
// --------------------------------------------------------------------------
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int wrapi(int v, int n) { int r = v % n; return r < 0 ? r + n : r; }
// Map an extended x index onto a real column: periodic wrap (run_mode >= 1) or
// clamped replication (walls), which is what the reference's ghost padding
// (YM_FCT.cu:7-28) and clamped block-halo loads (FCT_Sweeps.cu:52,73,78) produce.
__device__ __forceinline__ int map_x(int x, const KParams& p) {
    return p.periodic_x ? wrapi(x, p.nx) : clampi(x, 0, p.nx - 1);
}
// neighbour indices of the EM / source stencils
__device__ __forceinline__ int xm_src(int x, const KParams& p) {   // YM_Fluid.cu:38
    return p.periodic_x ? (x + p.nx - 1) % p.nx : (x > 0 ? x - 1 : 0);
}
__device__ __forceinline__ int xp_src(int x, const KParams& p) {   // YM_Physics.cu:148
    return p.periodic_x ? (x + 1) % p.nx : (x + 1 < p.nx ? x + 1 : p.nx - 1);
}
__device__ __forceinline__ int zm_per(int z, int nz) { return z == 0 ? nz - 1 : z - 1; }
__device__ __forceinline__ int zp_per(int z, int nz) { return z == nz - 1 ? 0 : z + 1; }

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), each with its own L2.  Remap the hardware block id so that XCD k
// walks the contiguous logical range [k*M, (k+1)*M), M = nb8/8: tiles that are
// adjacent in logical order (and share halos / field reads) then run on the
// same XCD at about the same time.  nb8 = grid size rounded up to 8; logical
// ids >= the real tile count must exit.
__device__ __forceinline__ int xcd_logical(int b, int nb8) { return (b & 7) * (nb8 >> 3) + (b >> 3); }
__host__ __device__ __forceinline__ int round8(int n) { return (n + 7) & ~7; }
// Logical tile order of a step kernel, forward or reversed (p.rev).  The
// step alternates the direction from one kernel to the next, so a kernel
// starts on the rows its predecessor touched last, which are still in the
// 256 MB memory-side cache.
__device__ __forceinline__ int tile_order(int lt, int total, int rev) { return rev ? total - 1 - lt : lt; }

// Element i of an array through a 32-bit byte offset (every array is below
// 4 GiB): the load can then take the array base from scalar registers and one
// vector offset shared by all arrays of a cell.
template <typename T>
__device__ __forceinline__ T ld32(const T* a, unsigned i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(a) + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st32(T* a, unsigned i, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(a) + i * (unsigned)sizeof(T)) = v;
}
// The same store marked non-temporal (streaming): the line is not kept in
// the caches for reuse, which suits outputs that no kernel reads soon.
template <typename T>
__device__ __forceinline__ void st32_nt(T* a, unsigned i, T v) {
    __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(a) + i * (unsigned)sizeof(T)));
}

// Raw buffer stores (non-temporal).  The buffer's range check drops a store
// whose byte offset is at or beyond the buffer's size, so lanes that must not
// write take offset kNoStore instead of a branch around the store: a store
// under a lane-divergent branch makes the compiler's wait counting assume the
// branch was not taken, and every later wait for an earlier load then also
// waits for the store to complete.  The config admits arrays of at most
// kNoStore bytes (NX*NZ*8 <= 0xFFFFFFF0), so kNoStore is always out of range.
constexpr unsigned kNoStore = 0xFFFFFFF0u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ void buf_st_nt(double v, __amdgpu_buffer_rsrc_t r, unsigned off) {
    typedef unsigned U2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, off, 0, 2);
}
__device__ __forceinline__ void buf_st_nt(float v, __amdgpu_buffer_rsrc_t r, unsigned off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 2);
}
// the same without the non-temporal hint (outputs read again soon)
__device__ __forceinline__ void buf_st(double v, __amdgpu_buffer_rsrc_t r, unsigned off) {
    typedef unsigned U2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, off, 0, 0);
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, unsigned off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}

// Whole-wavefront lane shifts (DPP wave_shr:1 / wave_shl:1 of GFX9):
// lane_from_left(v) in lane l is v of lane l-1, lane_from_right(v) is v of
// lane l+1; lane 0 (from the left) and lane 63 (from the right) receive 0.
// Every lane of the wavefront must be active.
__device__ __forceinline__ int dpp_shr1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_shl1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ float lane_from_left(float v) { return __int_as_float(dpp_shr1(__float_as_int(v))); }
__device__ __forceinline__ float lane_from_right(float v) { return __int_as_float(dpp_shl1(__float_as_int(v))); }
__device__ __forceinline__ double lane_shift64(double v, bool left) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
    const int rlo = left ? dpp_shr1(lo) : dpp_shl1(lo);
    const int rhi = left ? dpp_shr1(hi) : dpp_shl1(hi);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)rhi << 32) | (unsigned)rlo));
}
__device__ __forceinline__ double lane_from_left(double v) { return lane_shift64(v, true); }
__device__ __forceinline__ double lane_from_right(double v) { return lane_shift64(v, false); }

// Explicit selects: deterministic sign of zero (NaN propagates through the
// first operand of kmax(a, floor)).
template <typename T> __device__ __forceinline__ T kmin(T a, T b) { return (b < a) ? b : a; }
template <typename T> __device__ __forceinline__ T kmax(T a, T b) { return (a < b) ? b : a; }

// --------------------------------------------------------------------------
// FCT coefficients and limiter (FCT_DeviceMath.cuh:38-63, 108-125)
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void fct_nu_mu(T eps, T& nu, T& mu) {
    T scrh = kmin(T(1.0 / 6.0), fabs(eps));
    scrh = T(1.0 / 3.0) * scrh * scrh;
    nu = T(1.0 / 6.0) + T(1.0 / 3.0) * (eps * eps - scrh * scrh);
    mu = T(0.25) - T(0.5) * nu;
}

template <typename T>
__device__ __forceinline__ T fct_limit(T flx, T dm1, T d0, T dp1) {
    T fsgn = copysign(T(1.0), d0);
    T term = fsgn * dm1;
    T terp = fsgn * dp1;
    T fa = fabs(flx);
    // the two inner minima as hardware min (fmin: v_min_f64 / v_min_f32, the
    // reference's CUDA min): they differ from the select form only in the
    // sign of a zero result (and in dropping NaN, as CUDA's min does), and the
    // outer max with +0 returns +0 for either zero
    T lim = kmax(T(0.0), fmin(term, fmin(fa, terp)));
    return fsgn * lim;
}

// --------------------------------------------------------------------------
// Sources evaluated from a state (the reference stores them between steps;
// they are pure functions of the end-of-step state, main_ym.cu:450-455).
// --------------------------------------------------------------------------
// kernel_ym_lorentz (YM_Fluid.cu:32-84): x-component at real cell (x, z).
template <typename T>
__device__ __forceinline__ T lorentz_px(const State<T>& S, int sp, int x, int z, const KParams& p) {
    const int nx = p.nx;
    const int b = S_NA + sp * F_COUNT;
    const int xm = xm_src(x, p), zm = zm_per(z, p.nz);
    const unsigned i = (unsigned)x + (unsigned)z * (unsigned)nx;
    const unsigned ixm = (unsigned)xm + (unsigned)z * (unsigned)nx;
    const unsigned izm = (unsigned)x + (unsigned)zm * (unsigned)nx;
    const unsigned ixmzm = (unsigned)xm + (unsigned)zm * (unsigned)nx;
    T n_safe = kmax(ld32(S.a[b + F_N], i), T(0.05));
    T vz = ld32(S.a[b + F_PZ], i) / n_safe;
    T acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const T* Ex = S.a[S_EX1 + c];
        const T* By = S.a[S_BY1 + c];
        T Exn = T(0.5) * (ld32(Ex, i) + ld32(Ex, ixm));
        T Byn = T(0.25) * (ld32(By, i) + ld32(By, ixm) + ld32(By, izm) + ld32(By, ixmzm));
        acc[c] = ld32(S.a[b + F_Q1 + c], i) * (Exn + vz * Byn);
    }
    T s = -(acc[0] + acc[1] + acc[2]);
    if (p.warm_T > 0.0) {
        const T* n = S.a[b + F_N];
        const int xp = xp_src(x, p);
        T dndx = (ld32(n, (unsigned)xp + (unsigned)z * (unsigned)nx) - ld32(n, ixm)) / (T(2.0) * T(p.dx));
        s -= T(p.warm_T) * dndx;
    }
    return s;
}

// z-component of kernel_ym_lorentz (YM_Fluid.cu:32-84) from values: density,
// x momentum and charges of the cell, Ez at (x, z) and (x, z-1), By at
// (x, z), (x-1, z), (x, z-1), (x-1, z-1).  Cold part only.
template <typename T>
__device__ __forceinline__ T lorentz_pz_u(T vx, const T (&q)[3], const T (&ez)[3], const T (&ezm)[3],
                                          const T (&by)[3], const T (&byxm)[3], const T (&byzm)[3],
                                          const T (&byxmzm)[3]) {
    T acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T Ezn = T(0.5) * (ez[c] + ezm[c]);
        T Byn = T(0.25) * (by[c] + byxm[c] + byzm[c] + byxmzm[c]);
        acc[c] = q[c] * (Ezn - vx * Byn);
    }
    return -(acc[0] + acc[1] + acc[2]);
}
template <typename T>
__device__ __forceinline__ T lorentz_pz_v(T n, T px, const T (&q)[3], const T (&ez)[3], const T (&ezm)[3],
                                          const T (&by)[3], const T (&byxm)[3], const T (&byzm)[3],
                                          const T (&byxmzm)[3]) {
    return lorentz_pz_u(px / kmax(n, T(0.05)), q, ez, ezm, by, byxm, byzm, byxmzm);
}

// Warm-closure pressure term of the z source: -T dn/dz (centred, periodic z).
template <typename T>
__device__ __forceinline__ T warm_pz(const State<T>& S, int sp, int x, int z, const KParams& p) {
    const T* n = S.a[S_NA + sp * F_COUNT + F_N];
    const int zp = zp_per(z, p.nz), zm = zm_per(z, p.nz);
    T dndz = (n[(size_t)x + (size_t)zp * p.nx] - n[(size_t)x + (size_t)zm * p.nx]) / (T(2.0) * T(p.dz));
    return T(p.warm_T) * dndz;
}

// z-component of kernel_ym_lorentz at real cell (x, z).
template <typename T>
__device__ __forceinline__ T lorentz_pz(const State<T>& S, int sp, int x, int z, const KParams& p) {
    const int nx = p.nx;
    const int b = S_NA + sp * F_COUNT;
    const int xm = xm_src(x, p), zm = zm_per(z, p.nz);
    const unsigned i = (unsigned)x + (unsigned)z * (unsigned)nx;
    const unsigned ixm = (unsigned)xm + (unsigned)z * (unsigned)nx;
    const unsigned izm = (unsigned)x + (unsigned)zm * (unsigned)nx;
    const unsigned ixmzm = (unsigned)xm + (unsigned)zm * (unsigned)nx;
    T q[3], ez[3], ezm[3], by[3], byxm[3], byzm[3], byxmzm[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        q[c] = ld32(S.a[b + F_Q1 + c], i);
        ez[c] = ld32(S.a[S_EZ1 + c], i);
        ezm[c] = ld32(S.a[S_EZ1 + c], izm);
        by[c] = ld32(S.a[S_BY1 + c], i);
        byxm[c] = ld32(S.a[S_BY1 + c], ixm);
        byzm[c] = ld32(S.a[S_BY1 + c], izm);
        byxmzm[c] = ld32(S.a[S_BY1 + c], ixmzm);
    }
    T s = lorentz_pz_v(ld32(S.a[b + F_N], i), ld32(S.a[b + F_PX], i), q, ez, ezm, by, byxm, byzm, byxmzm);
    if (p.warm_T > 0.0) s -= warm_pz(S, sp, x, z, p);
    return s;
}

// kernel_ym_precession (YM_Fluid.cu:86-108) from values: density, z momentum
// and charges of the cell, Az at (x, z) and (x, z-1); out[0..2] = sQ^a.
template <typename T>
__device__ __forceinline__ void precession_u(T vz, const T (&q)[3], const T (&az)[3], const T (&azm)[3], T alpha,
                                             T out[3]) {
    T a1 = T(0.5) * (az[0] + azm[0]);
    T a2 = T(0.5) * (az[1] + azm[1]);
    T a3 = T(0.5) * (az[2] + azm[2]);
    out[0] = alpha * vz * (q[1] * a3 - q[2] * a2);
    out[1] = alpha * vz * (q[2] * a1 - q[0] * a3);
    out[2] = alpha * vz * (q[0] * a2 - q[1] * a1);
}
template <typename T>
__device__ __forceinline__ void precession_v(T n, T pz, const T (&q)[3], const T (&az)[3], const T (&azm)[3],
                                             T alpha, T out[3]) {
    precession_u(pz / kmax(n, T(0.05)), q, az, azm, alpha, out);
}

// kernel_ym_precession at real cell (x, z).
template <typename T>
__device__ __forceinline__ void precession_at(const State<T>& S, int sp, int x, int z,
                                              const KParams& p, T out[3]) {
    const int nx = p.nx;
    const int b = S_NA + sp * F_COUNT;
    const int zm = zm_per(z, p.nz);
    const unsigned i = (unsigned)x + (unsigned)z * (unsigned)nx;
    const unsigned izm = (unsigned)x + (unsigned)zm * (unsigned)nx;
    T q[3], az[3], azm[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        q[c] = ld32(S.a[b + F_Q1 + c], i);
        az[c] = ld32(S.a[S_AZ1 + c], i);
        azm[c] = ld32(S.a[S_AZ1 + c], izm);
    }
    precession_v(ld32(S.a[b + F_N], i), ld32(S.a[b + F_PZ], i), q, az, azm, T(p.alpha), out);
}

// kernel_ym_currents (YM_Physics.cu:9-60) at real cell (x, z) from fluid F.
template <typename T>
__device__ __forceinline__ void currents_at(const State<T>& F, int x, int z, const KParams& p,
                                            T Jx[3], T Jz[3]) {
    const int nx = p.nx;
    const int xp = xp_src(x, p), zp = zp_per(z, p.nz);
    const unsigned i = (unsigned)x + (unsigned)z * (unsigned)nx;
    const unsigned ixp = (unsigned)xp + (unsigned)z * (unsigned)nx;
    const unsigned izp = (unsigned)x + (unsigned)zp * (unsigned)nx;
    T vxr[2], vzt[2], nQx[2][3], nQz[2][3];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int b = S_NA + s * F_COUNT;
        const T* n = F.a[b + F_N];
        const T* px = F.a[b + F_PX];
        const T* pz = F.a[b + F_PZ];
        T n0 = ld32(n, i), np = ld32(n, ixp), nzp = ld32(n, izp);
        vxr[s] = T(0.5) * (ld32(px, i) / kmax(n0, T(1e-10)) + ld32(px, ixp) / kmax(np, T(1e-10)));
        vzt[s] = T(0.5) * (ld32(pz, i) / kmax(n0, T(1e-10)) + ld32(pz, izp) / kmax(nzp, T(1e-10)));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T* Q = F.a[b + F_Q1 + c];
            T q0 = ld32(Q, i);
            nQx[s][c] = T(0.5) * (n0 * q0 + np * ld32(Q, ixp));
            nQz[s][c] = T(0.5) * (n0 * q0 + nzp * ld32(Q, izp));
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Jx[c] = -(nQx[0][c] * vxr[0] + nQx[1][c] * vxr[1]);
        Jz[c] = -(nQz[0][c] * vzt[0] + nQz[1][c] * vzt[1]);
    }
}

// --------------------------------------------------------------------------
// EMHD (emhdgpu2d) state slots: fields, then fluid species at 12 and 18.
// --------------------------------------------------------------------------
enum EmSlot { E_EX = 0, E_EY, E_EZ, E_BX, E_BY, E_BZ, E_AX, E_AY, E_AZ };
enum EmVar { EV_N = 0, EV_PX, EV_PY, EV_PZ, EV_COUNT };
__host__ __device__ __forceinline__ int em_base(int sp) { return 12 + 6 * sp; }

// Field values around a cell that the EMHD Lorentz force reads from one row:
// (x, z) and (x-1, z).  The force also needs Ez, Bx, By, By(x-1) of row z-1.
template <typename T>
struct EmRow {
    T ex, exm, ey, ez, bx, bz, bzxm, by, byxm;
};

template <typename T>
__device__ __forceinline__ void em_load_row(const State<T>& S, unsigned i, unsigned ixm, EmRow<T>& r) {
    r.ex = ld32(S.a[E_EX], i); r.exm = ld32(S.a[E_EX], ixm);
    r.ey = ld32(S.a[E_EY], i); r.ez = ld32(S.a[E_EZ], i);
    r.bx = ld32(S.a[E_BX], i);
    r.bz = ld32(S.a[E_BZ], i); r.bzxm = ld32(S.a[E_BZ], ixm);
    r.by = ld32(S.a[E_BY], i); r.byxm = ld32(S.a[E_BY], ixm);
}

// kernel_compute_lorentz (emhdgpu2d/main_coupled.cu:112-136) from values:
// row z (c), row z-1 (m), the fluid cell (n, px, py, pz) and the charge q
// (the reference fluid has q = -1: "-1.0 * (...)").
template <typename T>
__device__ __forceinline__ void em_lorentz_v(const EmRow<T>& c, const EmRow<T>& m, T n, T px, T py, T pz, T q,
                                             T out[3]) {
    const T Ex_node = T(0.5) * (c.ex + c.exm);
    const T Ey_node = c.ey;
    const T Ez_node = T(0.5) * (c.ez + m.ez);
    const T Bx_node = T(0.5) * (c.bx + m.bx);
    const T Bz_node = T(0.5) * (c.bz + c.bzxm);
    const T By_node = T(0.25) * (c.by + c.byxm + m.by + m.byxm);
    const T n_safe = kmax(n, T(0.05));
    const T vx = px / n_safe;
    const T vy = py / n_safe;
    const T vz = pz / n_safe;
    out[0] = q * (Ex_node + (vy * Bz_node - vz * By_node));
    out[1] = q * (Ey_node + (vz * Bx_node - vx * Bz_node));
    out[2] = q * (Ez_node + (vx * By_node - vy * Bx_node));
}

// kernel_compute_lorentz at (x, z) for species sp.
template <typename T>
__device__ __forceinline__ void em_lorentz(const State<T>& S, int sp, int x, int z, const KParams& p,
                                           T out[3]) {
    const int nx = p.nx;
    const int xm = xm_src(x, p), zm = zm_per(z, p.nz);
    const size_t i = (size_t)x + (size_t)z * nx;
    const size_t ixm = (size_t)xm + (size_t)z * nx;
    const size_t izm = (size_t)x + (size_t)zm * nx;
    const size_t ixmzm = (size_t)xm + (size_t)zm * nx;
    EmRow<T> c, m;
    em_load_row(S, i, ixm, c);
    em_load_row(S, izm, ixmzm, m);
    const int b = em_base(sp);
    em_lorentz_v(c, m, S.a[b + EV_N][i], S.a[b + EV_PX][i], S.a[b + EV_PY][i], S.a[b + EV_PZ][i],
                 T(sp ? p.q1 : p.q0), out);
}

// --------------------------------------------------------------------------
// Model policies of the FCT sweeps: number of advected variables, their
// state slots, which variables carry a source in each sweep, and the
// source values (evaluated from the step-start state).
// --------------------------------------------------------------------------
struct YMModel {                    // ymgpu2d: n, px, pz, Q1..3 per beam
    static constexpr int NV = F_COUNT;
    static constexpr int IPX = F_PX, IPZ = F_PZ;
    static constexpr int NXS = 1;   // x sweep: px only (YM_FCT.cu:37-43)
    __host__ __device__ static constexpr int base(int sp) { return S_NA + sp * F_COUNT; }
    __host__ __device__ static constexpr int xslot(int v) { return v == F_PX ? 0 : -1; }
    template <typename T>
    __device__ static void x_src(const State<T>& S, int sp, int x, int z, const KParams& p, T (&s)[NXS]) {
        s[0] = lorentz_px(S, sp, x, z, p);
    }
    // z sweep: pz and Q^a (YM_FCT.cu:119-125)
    template <typename T>
    __device__ static void z_src(const State<T>& S, int sp, int x, int z, const KParams& p, T (&s)[NV]) {
        s[F_N] = T(0);
        s[F_PX] = T(0);
        s[F_PZ] = lorentz_pz(S, sp, x, z, p);
        T q[3];
        precession_at(S, sp, x, z, p, q);
        s[F_Q1] = q[0]; s[F_Q2] = q[1]; s[F_Q3] = q[2];
    }
    // Row-marching form of the z sweep's inputs: the field values of one row
    // that the sources read at (x, z) / (x-1, z); row z-1 comes from the
    // previous step of the march.
    template <typename T>
    struct Row {
        T ez[3], by[3], byxm[3], az[3];
    };
    template <typename T>
    __device__ static void load_row(const State<T>& S, unsigned i, unsigned ixm, Row<T>& r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            r.ez[c] = ld32(S.a[S_EZ1 + c], i);
            r.by[c] = ld32(S.a[S_BY1 + c], i);
            r.byxm[c] = ld32(S.a[S_BY1 + c], ixm);
            r.az[c] = ld32(S.a[S_AZ1 + c], i);
        }
    }
    template <typename T>
    __device__ static void sel_row(bool take, const Row<T>& a, Row<T>& r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            r.ez[c] = take ? a.ez[c] : r.ez[c];
            r.by[c] = take ? a.by[c] : r.by[c];
            r.byxm[c] = take ? a.byxm[c] : r.byxm[c];
            r.az[c] = take ? a.az[c] : r.az[c];
        }
    }
    // u (stale z velocity) and the z sources from the cell's fluid F (the
    // species' NV state values) and rows z (c) and z-1 (m).
    template <typename T>
    __device__ static void z_src_rows(const T (&F)[NV], const Row<T>& c, const Row<T>& m, const State<T>& S,
                                      int sp, int x, int z, const KParams& p, T& u, T (&s)[NV]) {
        const T n_safe = kmax(F[F_N], T(0.05));
        u = F[F_PZ] / n_safe;
        const T q[3] = {F[F_Q1], F[F_Q2], F[F_Q3]};
        s[F_N] = T(0);
        s[F_PX] = T(0);
        s[F_PZ] = lorentz_pz_v(F[F_N], F[F_PX], q, c.ez, m.ez, c.by, c.byxm, m.by, m.byxm);
        if (p.warm_T > 0.0) s[F_PZ] -= warm_pz(S, sp, x, z, p);
        T o[3];
        precession_v(F[F_N], F[F_PZ], q, c.az, m.az, T(p.alpha), o);
        s[F_Q1] = o[0]; s[F_Q2] = o[1]; s[F_Q3] = o[2];
    }
    // Variable groups of the z sweep (one wave each): G = 0 -> n, px, pz
    // (source on pz from the Lorentz force), G = 1 -> Q1..Q3 (precession).
    static constexpr int NG = 3;
    template <int G, typename T>
    __device__ static void load_row_g(const State<T>& S, unsigned i, unsigned ixm, Row<T>& r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (G == 0) {
                r.ez[c] = ld32(S.a[S_EZ1 + c], i);
                r.by[c] = ld32(S.a[S_BY1 + c], i);
                r.byxm[c] = ld32(S.a[S_BY1 + c], ixm);
            } else {
                r.az[c] = ld32(S.a[S_AZ1 + c], i);
            }
        }
    }
    template <int G, typename T>
    __device__ static void sel_row_g(bool take, const Row<T>& a, Row<T>& r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (G == 0) {
                r.ez[c] = take ? a.ez[c] : r.ez[c];
                r.by[c] = take ? a.by[c] : r.by[c];
                r.byxm[c] = take ? a.byxm[c] : r.byxm[c];
            } else {
                r.az[c] = take ? a.az[c] : r.az[c];
            }
        }
    }
    template <int G, typename T>
    __device__ static void z_src_rows_g(const T (&F)[NV], const Row<T>& c, const Row<T>& m, const State<T>& S,
                                        int sp, int x, int z, const KParams& p, T& u, T (&s)[NG]) {
        const T n_safe = kmax(F[F_N], T(0.05));
        u = F[F_PZ] / n_safe;
        const T q[3] = {F[F_Q1], F[F_Q2], F[F_Q3]};
        if (G == 0) {
            s[0] = T(0);
            s[1] = T(0);
            s[2] = lorentz_pz_v(F[F_N], F[F_PX], q, c.ez, m.ez, c.by, c.byxm, m.by, m.byxm);
            if (p.warm_T > 0.0) s[2] -= warm_pz(S, sp, x, z, p);
        } else {
            precession_v(F[F_N], F[F_PZ], q, c.az, m.az, T(p.alpha), s);
        }
    }
};

struct EMModel {                    // emhdgpu2d: n, px, py, pz; sources in both sweeps
    static constexpr int NV = EV_COUNT;
    static constexpr int IPX = EV_PX, IPZ = EV_PZ;
    static constexpr int NXS = 3;
    __host__ __device__ static constexpr int base(int sp) { return 12 + 6 * sp; }
    __host__ __device__ static constexpr int xslot(int v) { return v == EV_N ? -1 : v - 1; }
    template <typename T>
    __device__ static void x_src(const State<T>& S, int sp, int x, int z, const KParams& p, T (&s)[NXS]) {
        em_lorentz(S, sp, x, z, p, s);
    }
    template <typename T>
    __device__ static void z_src(const State<T>& S, int sp, int x, int z, const KParams& p, T (&s)[NV]) {
        T f[3];
        em_lorentz(S, sp, x, z, p, f);
        s[EV_N] = T(0); s[EV_PX] = f[0]; s[EV_PY] = f[1]; s[EV_PZ] = f[2];
    }
    template <typename T>
    using Row = EmRow<T>;
    template <typename T>
    __device__ static void load_row(const State<T>& S, unsigned i, unsigned ixm, Row<T>& r) {
        em_load_row(S, i, ixm, r);
    }
    template <typename T>
    __device__ static void sel_row(bool take, const Row<T>& a, Row<T>& r) {
        r.ex = take ? a.ex : r.ex; r.exm = take ? a.exm : r.exm;
        r.ey = take ? a.ey : r.ey; r.ez = take ? a.ez : r.ez;
        r.bx = take ? a.bx : r.bx;
        r.bz = take ? a.bz : r.bz; r.bzxm = take ? a.bzxm : r.bzxm;
        r.by = take ? a.by : r.by; r.byxm = take ? a.byxm : r.byxm;
    }
    template <typename T>
    __device__ static void z_src_rows(const T (&F)[NV], const Row<T>& c, const Row<T>& m, const State<T>&, int sp,
                                      int, int, const KParams& p, T& u, T (&s)[NV]) {
        const T n_safe = kmax(F[EV_N], T(0.05));
        u = F[EV_PZ] / n_safe;
        T f[3];
        em_lorentz_v(c, m, F[EV_N], F[EV_PX], F[EV_PY], F[EV_PZ], T(sp ? p.q1 : p.q0), f);
        s[EV_N] = T(0); s[EV_PX] = f[0]; s[EV_PY] = f[1]; s[EV_PZ] = f[2];
    }
    // variable groups: G = 0 -> n, px; G = 1 -> py, pz (both need the full
    // Lorentz force, which is cheap)
    static constexpr int NG = 2;
    template <int G, typename T>
    __device__ static void load_row_g(const State<T>& S, unsigned i, unsigned ixm, Row<T>& r) {
        em_load_row(S, i, ixm, r);
    }
    template <int G, typename T>
    __device__ static void sel_row_g(bool take, const Row<T>& a, Row<T>& r) {
        sel_row(take, a, r);
    }
    template <int G, typename T>
    __device__ static void z_src_rows_g(const T (&F)[NV], const Row<T>& c, const Row<T>& m, const State<T>&,
                                        int sp, int, int, const KParams& p, T& u, T (&s)[NG]) {
        const T n_safe = kmax(F[EV_N], T(0.05));
        u = F[EV_PZ] / n_safe;
        T f[3];
        em_lorentz_v(c, m, F[EV_N], F[EV_PX], F[EV_PY], F[EV_PZ], T(sp ? p.q1 : p.q0), f);
        if (G == 0) { s[0] = T(0); s[1] = f[0]; }
        else { s[0] = f[1]; s[1] = f[2]; }
    }
};

}  // namespace khs
