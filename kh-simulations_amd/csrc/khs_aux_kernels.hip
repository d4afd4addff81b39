// khs_aux_kernels.hip — filters, diagnostics and derived fields on gfx950.
//
//   k_cell_filters  sponge damping / xi_cut zeroing of the colour-2/3 set
//                   (YM_Physics.cu:161-210), one pass, per-column action table.
//   k_col_coef2     DFT coefficients of whole z columns: 64 columns per block
//   k_col_apply     (coalesced), z split over 8 row groups, modes in registers;
//                   then subtraction (reference band filter,
//                   YM_Physics.cu:349-584; z-mean removal :272-327) or
//                   projection (keep |kz| <= K extension).
//   k_hyperdiff     4th-order z hyperdiffusion (YM_Physics.cu:228-259) as an
//                   in-place column march with a register window of the old
//                   values, i.e. Jacobi semantics without a scratch copy.
//   k_energy_*      FP64 compensated (Neumaier) energy reduction: per-lane sums,
//   k_gauss_*       LDS tree of (sum, error) pairs merged with TwoSum, fixed
//                   launch shape so the result is deterministic.
//   k_derived       velocities / sources / currents / energy density of a state.
#include "khs_device.h"
#include "khs_launch.h"

namespace khs {

// --------------------------------------------------------------------------
// sponge / xi_cut
// --------------------------------------------------------------------------
// One thread per column and CF_ROWS rows: the column's action is read once,
// columns without one return at once (no per-element index arithmetic).  The
// last nzero arrays of the list are set to zero in every column (the
// colour-1 EM zeroing of suppress_kz0, which no other filter of the stage
// touches, in the same launch).
constexpr int CF_ROWS = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_cell_filters(ArrayList<T> arrs, int nzero, const double* decay,
                                                      const int* action, KParams p) {
    const int x = blockIdx.x * 256 + (int)threadIdx.x;
    if (x >= p.nx) return;
    const int z0 = blockIdx.y * CF_ROWS, z1 = min(p.nz, z0 + CF_ROWS);
    const int ncut = arrs.n - nzero;
    for (int k = ncut; k < arrs.n; ++k) {
        T* a = arrs.a[k] + x;
        for (int z = z0; z < z1; ++z) a[(size_t)z * p.nx] = T(0.0);
    }
    const int act = action ? action[x] : 0;
    if (act == 0) return;
    const T d = (act == 1) ? T(decay[x]) : T(0.0);
    for (int k = 0; k < ncut; ++k) {
        T* a = arrs.a[k] + x;
        if (act == 1) {
            for (int z = z0; z < z1; ++z) a[(size_t)z * p.nx] *= d;
        } else {
            for (int z = z0; z < z1; ++z) a[(size_t)z * p.nx] = T(0.0);
        }
    }
}

template <typename T>
void launch_cell_filters(const ArrayList<T>& arrs, const double* decay, const int* action,
                         const KParams& p, hipStream_t st, int nzero) {
    hipLaunchKernelGGL(k_cell_filters<T>, dim3((p.nx + 255) / 256, (p.nz + CF_ROWS - 1) / CF_ROWS), dim3(256), 0, st,
                       arrs, nzero, decay, action, p);
}

// --------------------------------------------------------------------------
// z-mean removal with the charge cut and the colour-1 zeroing in one pass
// (the filter set of suppress_kz0 runs without bands, keep or hyperdiffusion:
// main_ym.cu:365-447 in that case is sponge / xi_cut on Q2, Q3, z-mean of 15
// arrays, colour-1 EM set to zero).  A workgroup owns a 64-column strip of
// one array: arrays [0, nmean) lose their column mean (the last nact of them
// first take the column action: 1 = times decay[x], 2 = zero; per column, so
// it commutes with the column mean), arrays [nmean, n) are set to zero.  The
// column sum is FP64 over ZM_G row groups (each with four running sums)
// added in group order; the mean is sum * w (w = 1 / nz, as in the column
// DFT filter's mode-0 weight).
// --------------------------------------------------------------------------
constexpr int ZM_G = 16;     // row groups per workgroup of k_zmean_cut

// One 64-column strip of array k (the workgroup: 64 lanes x ZM_G row groups).
template <typename T>
__device__ __forceinline__ void zmean_cut_strip(const ArrayList<T>& arrs, int nmean, int nact, const double* decay,
                                                const int* action, double w, const KParams& p, int strip, int k) {
    // rows per thread held in registers (nz <= 16 R): 64 FP32 / 32 FP64
    // values, within the 128 VGPRs a 1024-thread workgroup allows
    constexpr int R = sizeof(T) == 4 ? 64 : 32;
    __shared__ double red[ZM_G][64];
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int x = strip * 64 + lane;
    const int nx = p.nx, nz = p.nz;
    const int rows = (nz + ZM_G - 1) / ZM_G;
    const int z0 = g * rows, z1 = min(nz, z0 + rows);
    T* v = arrs.a[k];
    if (k >= nmean) {
        if (x < nx)
#pragma unroll 4
            for (int z = z0; z < z1; ++z) v[(size_t)x + (size_t)z * nx] = T(0.0);
        return;                                        // whole workgroups leave together
    }
    int act = 0;
    T d = T(1.0);
    if (x < nx && k >= nmean - nact && action) {
        act = action[x];
        if (act == 1) d = T(decay[x]);
    }
    auto val = [&](size_t i) {                         // the value after the column action
        T a = v[i];
        if (act == 1) a *= d;
        else if (act == 2) a = T(0.0);
        return a;
    };
    // four running sums, added in order: rows z0 + 4j .. z0 + 4j + 3 go to
    // s0..s3, the rows after the last whole group of four to s0
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    auto column_mean = [&](double tsum) {
        red[g][lane] = tsum;
        __syncthreads();
        double tot = 0.0;
#pragma unroll
        for (int q = 0; q < ZM_G; ++q) tot += red[q][lane];
        return tot * w;
    };
    if (z1 - z0 <= R) {
        // one pass: the thread's rows stay in registers between the sum and
        // the subtraction (the same sums in the same order as below)
        const int cnt = z1 - z0;
        T c[R];
#pragma unroll
        for (int q = 0; q < R; ++q)
            c[q] = (x < nx && q < cnt) ? val((size_t)x + (size_t)(z0 + q) * nx) : T(0.0);
#pragma unroll
        for (int q = 0; q < R; q += 4) {
            if (q + 4 <= cnt) {
                s0 += (double)c[q]; s1 += (double)c[q + 1]; s2 += (double)c[q + 2]; s3 += (double)c[q + 3];
            } else {
#pragma unroll
                for (int r = q; r < q + 4; ++r)
                    if (r < cnt) s0 += (double)c[r];
            }
        }
        const double mean = column_mean((s0 + s1) + (s2 + s3));
        if (x < nx)
#pragma unroll
            for (int q = 0; q < R; ++q)
                if (q < cnt) v[(size_t)x + (size_t)(z0 + q) * nx] = (T)((double)c[q] - mean);
        return;
    }
    // taller columns: two passes (four loads in flight), the second read of
    // the strip mostly from the cache
    if (x < nx) {
        int z = z0;
        for (; z + 4 <= z1; z += 4) {
            const T a0 = val((size_t)x + (size_t)z * nx), a1 = val((size_t)x + (size_t)(z + 1) * nx);
            const T a2 = val((size_t)x + (size_t)(z + 2) * nx), a3 = val((size_t)x + (size_t)(z + 3) * nx);
            s0 += (double)a0; s1 += (double)a1; s2 += (double)a2; s3 += (double)a3;
        }
        for (; z < z1; ++z) s0 += (double)val((size_t)x + (size_t)z * nx);
    }
    const double mean = column_mean((s0 + s1) + (s2 + s3));
    if (x < nx)
#pragma unroll 4
        for (int z = z0; z < z1; ++z) {
            const size_t i = (size_t)x + (size_t)z * nx;
            v[i] = (T)((double)val(i) - mean);
        }
}

template <typename T>
__global__ __launch_bounds__(64 * ZM_G) void k_zmean_cut(ArrayList<T> arrs, int nmean, int nact, const double* decay,
                                                         const int* action, double w, KParams p) {
    zmean_cut_strip<T>(arrs, nmean, nact, decay, action, w, p, blockIdx.x, blockIdx.y);
}

// The sponge / xi_cut column action on a 64-column strip of array k (the
// workgroup's 64 lanes x ZM_G row groups), the arithmetic of k_cell_filters.
template <typename T>
__device__ __forceinline__ void cell_filter_strip(const ArrayList<T>& arrs, const double* decay, const int* action,
                                                  const KParams& p, int strip, int k) {
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int x = strip * 64 + lane;
    if (x >= p.nx) return;
    const int act = action[x];
    if (act == 0) return;
    const int rows = (p.nz + ZM_G - 1) / ZM_G;
    const int z0 = g * rows, z1 = min(p.nz, z0 + rows);
    const T d = (act == 1) ? T(decay[x]) : T(0.0);
    T* a = arrs.a[k] + x;
    if (act == 1) {
        for (int z = z0; z < z1; ++z) a[(size_t)z * p.nx] *= d;
    } else {
        for (int z = z0; z < z1; ++z) a[(size_t)z * p.nx] = T(0.0);
    }
}

// Filter stage of packed runs: work item i of run r is strip i % nstrip of
// filtered array i / nstrip.
template <typename T>
__global__ __launch_bounds__(64 * ZM_G) void k_filters_packed(const PackedRun<T>* __restrict__ runs,
                                                              const int* __restrict__ begin, int n, int total) {
    const int t = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
    if (t >= total) return;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int r = __builtin_amdgcn_readfirstlane(lo);
    const PackedRun<T>& d = runs[r];
    const int i = t - begin[r];
    const int strip = i % d.nstrip, k = i / d.nstrip;
    // the run's parameters and array table are read in place (a local copy
    // of the table, indexed by k at run time, would live in scratch memory)
    const KParams& p = d.p;
    const ArrayList<T>& arrs = d.filt;
    if (d.fmode == 2)
        zmean_cut_strip<T>(arrs, d.nmean, d.nact, d.decay, d.action, 1.0 / (double)p.nz, p, strip, k);
    else if (d.fmode == 1)
        cell_filter_strip<T>(arrs, d.decay, d.action, p, strip, k);
}

template <typename T>
void launch_filters_packed(const PackedRun<T>* runs, const int* begin, int n, int total, hipStream_t st) {
    if (total > 0)
        hipLaunchKernelGGL(k_filters_packed<T>, dim3(total), dim3(64 * ZM_G), 0, st, runs, begin, n, total);
}
template void launch_filters_packed<float>(const PackedRun<float>*, const int*, int, int, hipStream_t);
template void launch_filters_packed<double>(const PackedRun<double>*, const int*, int, int, hipStream_t);

template <typename T>
void launch_zmean_cut(const ArrayList<T>& arrs, int nmean, int nact, const double* decay, const int* action,
                      const KParams& p, hipStream_t st) {
    if (arrs.n == 0) return;
    hipLaunchKernelGGL(k_zmean_cut<T>, dim3((p.nx + 63) / 64, arrs.n), dim3(64 * ZM_G), 0, st, arrs, nmean, nact,
                       decay, action, 1.0 / (double)p.nz, p);
}

// --------------------------------------------------------------------------
// column DFT filter
// --------------------------------------------------------------------------
constexpr int CZ = 64;   // z rows per apply block

// DFT coefficients of whole z columns, FP64 sums: a block is 64 columns x CG
// groups of rows; every group sums its rows for CM2 modes at a time, and the
// groups' partial sums are added in group order through LDS.  A column costs
// nz / CG dependent iterations, which is what narrow, tall grids need (a
// 256 x 1024 grid has only 4 column strips).
constexpr int CG = 8;      // z groups per block
constexpr int CM2 = 8;     // modes per pass

template <typename T, typename TW>
__global__ __launch_bounds__(64 * CG) void k_col_coef2(ArrayList<T> arrs, const TW* ctab, const TW* stab,
                                                       const TW* weight, int nmodes, double* coef, KParams p) {
    __shared__ double red[CG][CM2][2][64];
    const int lane = threadIdx.x & 63;
    const int g = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + lane;
    const int k = blockIdx.y;
    const int m0 = blockIdx.z * CM2;
    const int mc = min(CM2, nmodes - m0);
    const int rows = (p.nz + CG - 1) / CG;
    const int z0 = g * rows, z1 = min(p.nz, z0 + rows);
    double sc[CM2], ss[CM2];
#pragma unroll
    for (int j = 0; j < CM2; ++j) { sc[j] = 0.0; ss[j] = 0.0; }
    if (x < p.nx) {
        const T* v = arrs.a[k];
        for (int z = z0; z < z1; ++z) {
            const double val = (double)v[(size_t)x + (size_t)z * p.nx];
#pragma unroll
            for (int j = 0; j < CM2; ++j) {
                if (j < mc) {
                    sc[j] += val * (double)ctab[(size_t)(m0 + j) * p.nz + z];
                    ss[j] += val * (double)stab[(size_t)(m0 + j) * p.nz + z];
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < CM2; ++j) { red[g][j][0][lane] = sc[j]; red[g][j][1][lane] = ss[j]; }
    __syncthreads();
    if (g == 0 && x < p.nx) {
#pragma unroll
        for (int j = 0; j < CM2; ++j) {
            if (j < mc) {
                double a = red[0][j][0][lane], b = red[0][j][1][lane];
                for (int q = 1; q < CG; ++q) { a += red[q][j][0][lane]; b += red[q][j][1][lane]; }
                const size_t base = ((size_t)k * nmodes + (m0 + j)) * 2;
                coef[(base + 0) * p.nx + x] = a * (double)weight[m0 + j];
                coef[(base + 1) * p.nx + x] = b * (double)weight[m0 + j];
            }
        }
    }
}

template <typename T, typename TW>
__global__ __launch_bounds__(64) void k_col_apply(ArrayList<T> arrs, const TW* ctab, const TW* stab,
                                                  int nmodes, int project, const double* coef, KParams p) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int k = blockIdx.y;
    const int zb = blockIdx.z * CZ;
    if (x >= p.nx) return;
    T* v = arrs.a[k];
    const int ze = min(zb + CZ, p.nz);
    for (int z = zb; z < ze; ++z) {
        double delta = 0.0;
        for (int m = 0; m < nmodes; ++m) {
            const size_t base = ((size_t)k * nmodes + m) * 2;
            const double A = coef[(base + 0) * p.nx + x];
            const double B = coef[(base + 1) * p.nx + x];
            delta += A * (double)ctab[(size_t)m * p.nz + z] + B * (double)stab[(size_t)m * p.nz + z];
        }
        const size_t i = (size_t)x + (size_t)z * p.nx;
        if (project) v[i] = (T)delta;
        else v[i] = (T)((double)v[i] - delta);
    }
}

template <typename T, typename TW>
void launch_column_modes(const ArrayList<T>& arrs, const TW* ctab, const TW* stab,
                         const TW* weight, int nmodes, int project, double* coef,
                         const KParams& p, hipStream_t st) {
    if (arrs.n == 0 || nmodes == 0) return;
    dim3 g1((p.nx + 63) / 64, arrs.n, (nmodes + CM2 - 1) / CM2);
    hipLaunchKernelGGL((k_col_coef2<T, TW>), g1, dim3(64 * CG), 0, st, arrs, ctab, stab, weight, nmodes, coef, p);
    dim3 g2((p.nx + 63) / 64, arrs.n, (p.nz + CZ - 1) / CZ);
    hipLaunchKernelGGL((k_col_apply<T, TW>), g2, dim3(64), 0, st, arrs, ctab, stab, nmodes, project, coef, p);
}

// --------------------------------------------------------------------------
// hyperdiffusion (requires nz >= 5)
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void k_hyperdiff(ArrayList<T> arrs, double mu_d, KParams p) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int k = blockIdx.y;
    if (x >= p.nx) return;
    T* v = arrs.a[k];
    const int nz = p.nz;
    const size_t nx = p.nx;
    const T mu = T(mu_d);
    auto at = [&](int z) { return v[(size_t)x + (size_t)z * nx]; };
    const T o0 = at(0), o1 = at(1), o2 = at(2);
    T zm2 = at(nz - 2), zm1 = at(nz - 1), c = o0, zp1 = o1, zp2 = o2;
    for (int z = 0; z < nz; ++z) {
        T d4 = zm2 - T(4.0) * zm1 + T(6.0) * c - T(4.0) * zp1 + zp2;
        T nw = c - mu * d4;
        // next old value (z+3), read before any write reaches it
        const int z3 = z + 3;
        T nxt = (z3 < nz) ? at(z3) : (z3 - nz == 0 ? o0 : (z3 - nz == 1 ? o1 : o2));
        v[(size_t)x + (size_t)z * nx] = nw;
        zm2 = zm1; zm1 = c; c = zp1; zp1 = zp2; zp2 = nxt;
    }
}

template <typename T>
void launch_hyperdiffuse(const ArrayList<T>& arrs, double mu, const KParams& p, hipStream_t st) {
    if (arrs.n == 0) return;
    dim3 g((p.nx + 63) / 64, arrs.n);
    hipLaunchKernelGGL(k_hyperdiff<T>, g, dim3(64), 0, st, arrs, mu, p);
}

// kernel_enforce_boundaries (YM_Fluid.cu:13-30; EMHD main_coupled.cu:48-63)
// on one state set, in place: density floor 0.05 of every fluid; with walls
// in x, px = 0 in the wall columns and the wall fields (YM: Ez^a; EMHD: Ey,
// Ez, Bx) = 0.  Cell-local.
template <typename T>
__global__ void k_enforce_boundaries(State<T> S, KParams p, int emhd) {
    const size_t n = (size_t)p.nx * p.nz;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % (size_t)p.nx);
        const bool wall = !p.periodic_x && (x == 0 || x == p.nx - 1);
        for (int s = 0; s < p.nsp; ++s) {
            const int b = emhd ? 12 + 6 * s : S_NA + s * F_COUNT;
            T* nn = S.a[b];
            if (nn[i] < T(0.05)) nn[i] = T(0.05);
            if (wall) S.a[b + 1][i] = T(0.0);                 // px (both models: slot b + 1)
        }
        if (wall) {
            if (emhd) {
                S.a[1][i] = T(0.0); S.a[2][i] = T(0.0); S.a[3][i] = T(0.0);   // Ey, Ez, Bx
            } else {
                S.a[S_EZ1][i] = T(0.0); S.a[S_EZ2][i] = T(0.0); S.a[S_EZ3][i] = T(0.0);
            }
        }
    }
}

template <typename T>
void launch_enforce_boundaries(const State<T>& S, const KParams& p, int emhd, hipStream_t st) {
    const size_t n = (size_t)p.nx * p.nz;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_enforce_boundaries<T>, dim3(blocks), dim3(256), 0, st, S, p, emhd);
}

// --------------------------------------------------------------------------
// compensated reductions
// --------------------------------------------------------------------------
// (CSum, two_sum, cadd, cmerge: khs_device.h)

template <int K>
__device__ void block_reduce_store(CSum (&acc)[K], double* partials) {
    __shared__ double ss[K][kRedThreads], sc[K][kRedThreads];
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k) { ss[k][t] = acc[k].s; sc[k][t] = acc[k].c; }
    __syncthreads();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if (t < off) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                CSum r = cmerge(CSum{ss[k][t], sc[k][t]}, CSum{ss[k][t + off], sc[k][t + off]});
                ss[k][t] = r.s; sc[k][t] = r.c;
            }
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            partials[((size_t)blockIdx.x * K + k) * 2 + 0] = ss[k][0];
            partials[((size_t)blockIdx.x * K + k) * 2 + 1] = sc[k][0];
        }
    }
}

template <int K>
__global__ __launch_bounds__(kRedThreads) void k_reduce_final(const double* partials, int nblocks, double* out) {
    CSum acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = CSum{0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += kRedThreads)
#pragma unroll
        for (int k = 0; k < K; ++k)
            acc[k] = cmerge(acc[k], CSum{partials[((size_t)b * K + k) * 2], partials[((size_t)b * K + k) * 2 + 1]});
    __shared__ double ss[K][kRedThreads], sc[K][kRedThreads];
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k) { ss[k][t] = acc[k].s; sc[k][t] = acc[k].c; }
    __syncthreads();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if (t < off) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                CSum r = cmerge(CSum{ss[k][t], sc[k][t]}, CSum{ss[k][t + off], sc[k][t + off]});
                ss[k][t] = r.s; sc[k][t] = r.c;
            }
        }
        __syncthreads();
    }
    if (t == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = ss[k][0] + sc[k][0];
}

// per-cell energy density, FP64 (YM_Energy.cu:3-21)
template <typename T>
__device__ __forceinline__ double energy_cell(const State<T>& S, size_t i) {
    double nA = (double)S.a[S_NA][i], nB = (double)S.a[S_NB][i];
    nA = kmax(nA, 1e-10); nB = kmax(nB, 1e-10);
    double pxA = S.a[S_PXA][i], pzA = S.a[S_PZA][i], pxB = S.a[S_PXB][i], pzB = S.a[S_PZB][i];
    double keA = 0.5 * (pxA * pxA + pzA * pzA) / nA;
    double keB = 0.5 * (pxB * pxB + pzB * pzB) / nB;
    const int ids[9] = {S_EX1, S_EX2, S_EX3, S_EZ1, S_EZ2, S_EZ3, S_BY1, S_BY2, S_BY3};
    double e0 = S.a[ids[0]][i];
    double acc = e0 * e0;
#pragma unroll
    for (int k = 1; k < 9; ++k) { double f = S.a[ids[k]][i]; acc = acc + f * f; }
    double em = 0.5 * acc;
    return keA + keB + em;
}

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_energy(State<T> S, KParams p, double* partials) {
    const size_t N = (size_t)p.nx * p.nz;
    CSum acc[1] = {CSum{0.0, 0.0}};
    for (size_t i = blockIdx.x * (size_t)kRedThreads + threadIdx.x; i < N; i += (size_t)gridDim.x * kRedThreads)
        cadd(acc[0], energy_cell(S, i));
    block_reduce_store<1>(acc, partials);
}

// First stage for long lists of partial sums: block b merges the CSum pairs
// [4 kRedThreads b, 4 kRedThreads (b + 1)) into out[b].
__global__ __launch_bounds__(kRedThreads) void k_reduce_stage(const double* __restrict__ in, int n, double* out) {
    CSum acc[1] = {CSum{0.0, 0.0}};
    const int base = blockIdx.x * 4 * kRedThreads + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = base + k * kRedThreads;
        if (i < n) acc[0] = cmerge(acc[0], CSum{in[2 * (size_t)i], in[2 * (size_t)i + 1]});
    }
    block_reduce_store<1>(acc, out);
}

// the compensated sum of n partial sums (CSum pairs) into out[0]; lists
// longer than 4 kRedThreads go through one first stage into scratch (room
// for ceil(n / (4 kRedThreads)) pairs), in a fixed order
void launch_reduce_partials(const double* partials, int n, double* scratch, double* out, hipStream_t st) {
    if (n <= 4 * kRedThreads) {
        hipLaunchKernelGGL(k_reduce_final<1>, dim3(1), dim3(kRedThreads), 0, st, partials, n, out);
        return;
    }
    const int nb = (n + 4 * kRedThreads - 1) / (4 * kRedThreads);
    hipLaunchKernelGGL(k_reduce_stage, dim3(nb), dim3(kRedThreads), 0, st, partials, n, scratch);
    hipLaunchKernelGGL(k_reduce_final<1>, dim3(1), dim3(kRedThreads), 0, st, scratch, nb, out);
}

template <typename T>
void launch_energy(const State<T>& S, const KParams& p, double* partials, double* out, hipStream_t st) {
    size_t N = (size_t)p.nx * p.nz;
    int blocks = (int)std::min<size_t>((N + kRedThreads - 1) / kRedThreads, kRedBlocks);
    hipLaunchKernelGGL(k_energy<T>, dim3(blocks), dim3(kRedThreads), 0, st, S, p, partials);
    hipLaunchKernelGGL(k_reduce_final<1>, dim3(1), dim3(kRedThreads), 0, st, partials, blocks, out);
}

// Gauss residual, analysis/gauss_check.py:40-115, periodic differences in x and z.
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_gauss(State<T> S, KParams p, double* partials) {
    const int nx = p.nx, nz = p.nz;
    const size_t N = (size_t)nx * nz;
    CSum acc[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) acc[k] = CSum{0.0, 0.0};
    const double DX = p.dx, DZ = p.dz, al = p.alpha;
    for (size_t i = blockIdx.x * (size_t)kRedThreads + threadIdx.x; i < N; i += (size_t)gridDim.x * kRedThreads) {
        const int x = (int)(i % nx), z = (int)(i / nx);
        const size_t ixm = (size_t)((x + nx - 1) % nx) + (size_t)z * nx;
        const size_t izm = (size_t)x + (size_t)((z + nz - 1) % nz) * nx;
        double dE[3], Ezc[3], Az[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T* Ex = S.a[S_EX1 + c];
            const T* Ez = S.a[S_EZ1 + c];
            dE[c] = ((double)Ex[i] - (double)Ex[ixm]) / DX + ((double)Ez[i] - (double)Ez[izm]) / DZ;
            Ezc[c] = 0.5 * ((double)Ez[i] + (double)Ez[izm]);
            Az[c] = S.a[S_AZ1 + c][i];
        }
        double nab[3] = {al * (Az[1] * Ezc[2] - Az[2] * Ezc[1]),
                         al * (Az[2] * Ezc[0] - Az[0] * Ezc[2]),
                         al * (Az[0] * Ezc[1] - Az[1] * Ezc[0])};
        double nA = S.a[S_NA][i], nB = S.a[S_NB][i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double rho = (double)S.a[S_Q1A + c][i] * nA + (double)S.a[S_Q1B + c][i] * nB;
            double Gp = dE[c] + nab[c] - rho;
            double Gm = dE[c] + nab[c] + rho;
            cadd(acc[c], Gp * Gp);
            cadd(acc[3 + c], Gm * Gm);
        }
        double ae = fabs((double)S.a[S_EX1][i]);
        ae = ae + fabs((double)S.a[S_EX2][i]);
        ae = ae + fabs((double)S.a[S_EX3][i]);
        ae = ae + fabs((double)S.a[S_EZ1][i]);
        ae = ae + fabs((double)S.a[S_EZ2][i]);
        ae = ae + fabs((double)S.a[S_EZ3][i]);
        cadd(acc[6], ae * ae);
    }
    block_reduce_store<7>(acc, partials);
}

template <typename T>
void launch_gauss(const State<T>& S, const KParams& p, double* partials, double* out, hipStream_t st) {
    size_t N = (size_t)p.nx * p.nz;
    int blocks = (int)std::min<size_t>((N + kRedThreads - 1) / kRedThreads, kRedBlocks);
    hipLaunchKernelGGL(k_gauss<T>, dim3(blocks), dim3(kRedThreads), 0, st, S, p, partials);
    hipLaunchKernelGGL(k_reduce_final<7>, dim3(1), dim3(kRedThreads), 0, st, partials, blocks, out);
}

// --------------------------------------------------------------------------
// derived arrays
// --------------------------------------------------------------------------
template <typename T>
__global__ void k_derived(State<T> S, int id, KParams p, double* out) {
    const size_t N = (size_t)p.nx * p.nz;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
         i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % p.nx), z = (int)(i / p.nx);
        double r = 0.0;
        if (id >= 32 && id <= 37) {                       // Jx1..3, Jz1..3
            T Jx[3], Jz[3];
            currents_at(S, x, z, p, Jx, Jz);
            r = (id < 35) ? (double)Jx[id - 32] : (double)Jz[id - 35];
        } else if (id >= 38 && id <= 41) {                // uxA, uzA, uxB, uzB
            const int sp = (id - 38) / 2, comp = (id - 38) % 2;
            const int b = S_NA + sp * F_COUNT;
            T n_safe = kmax(S.a[b + F_N][i], T(0.05));
            r = (double)(S.a[b + (comp ? F_PZ : F_PX)][i] / n_safe);
        } else if (id >= 42 && id <= 45) {                // src px/pz of A, B
            const int sp = (id - 42) / 2, comp = (id - 42) % 2;
            r = comp ? (double)lorentz_pz(S, sp, x, z, p) : (double)lorentz_px(S, sp, x, z, p);
        } else if (id >= 46 && id <= 51) {                // sQ1..3 of A, B
            const int sp = (id - 46) / 3, c = (id - 46) % 3;
            T q[3];
            precession_at(S, sp, x, z, p, q);
            r = (double)q[c];
        } else if (id == 52) {
            r = energy_cell(S, i);
        }
        out[i] = r;
    }
}

template <typename T>
void launch_derived(const State<T>& S, int id, const KParams& p, double* out, hipStream_t st) {
    size_t N = (size_t)p.nx * p.nz;
    int blocks = (int)std::min<size_t>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_derived<T>, dim3(blocks), dim3(256), 0, st, S, id, p, out);
}


// --------------------------------------------------------------------------
// EMHD energy and derived arrays (emhdgpu2d/main_coupled.cu:99-136, 197-222)
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ double energy_cell_em(const State<T>& S, int nsp, size_t i) {
    double ke = 0.0;
    for (int s = 0; s < nsp; ++s) {
        const int b = em_base(s);
        double n_safe = kmax((double)S.a[b + EV_N][i], 1e-10);
        double px = S.a[b + EV_PX][i], py = S.a[b + EV_PY][i], pz = S.a[b + EV_PZ][i];
        double k = 0.5 * (px * px + py * py + pz * pz) / n_safe;
        ke = (s == 0) ? k : ke + k;
    }
    double ex = S.a[E_EX][i], ey = S.a[E_EY][i], ez = S.a[E_EZ][i];
    double bx = S.a[E_BX][i], by = S.a[E_BY][i], bz = S.a[E_BZ][i];
    double em = 0.5 * (ex * ex + ey * ey + ez * ez + bx * bx + by * by + bz * bz);
    return ke + em;
}

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_energy_em(State<T> S, KParams p, double* partials) {
    const size_t N = (size_t)p.nx * p.nz;
    CSum acc[1] = {CSum{0.0, 0.0}};
    for (size_t i = blockIdx.x * (size_t)kRedThreads + threadIdx.x; i < N; i += (size_t)gridDim.x * kRedThreads)
        cadd(acc[0], energy_cell_em(S, p.nsp, i));
    block_reduce_store<1>(acc, partials);
}

template <typename T>
void launch_energy_em(const State<T>& S, const KParams& p, double* partials, double* out, hipStream_t st) {
    size_t N = (size_t)p.nx * p.nz;
    int blocks = (int)std::min<size_t>((N + kRedThreads - 1) / kRedThreads, kRedBlocks);
    hipLaunchKernelGGL(k_energy_em<T>, dim3(blocks), dim3(kRedThreads), 0, st, S, p, partials);
    hipLaunchKernelGGL(k_reduce_final<1>, dim3(1), dim3(kRedThreads), 0, st, partials, blocks, out);
}

template <typename T>
__global__ void k_derived_em(State<T> S, int which, KParams p, double* out) {
    const size_t N = (size_t)p.nx * p.nz;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
         i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % p.nx), z = (int)(i / p.nx);
        double r = 0.0;
        if (which <= 2) {                                   // J (kernel_compute_currents)
            const int xp = xp_src(x, p), zp = zp_per(z, p.nz);
            const size_t ixp = (size_t)xp + (size_t)z * p.nx, izp = (size_t)x + (size_t)zp * p.nx;
            T j = T(0);
            for (int s = 0; s < p.nsp; ++s) {
                const int b = em_base(s);
                const T q = T(s ? p.q1 : p.q0), h = T((s ? p.q1 : p.q0) * 0.5);
                T v;
                if (which == 0) v = h * (S.a[b + EV_PX][i] + S.a[b + EV_PX][ixp]);
                else if (which == 1) v = q * S.a[b + EV_PY][i];
                else v = h * (S.a[b + EV_PZ][i] + S.a[b + EV_PZ][izp]);
                j = (s == 0) ? v : j + v;
            }
            r = (double)j;
        } else if (which <= 6) {                            // velocities
            const int s = (which - 3) / 2, comp = (which - 3) % 2;
            const int b = em_base(s);
            T n_safe = kmax(S.a[b + EV_N][i], T(0.05));
            r = (double)(S.a[b + (comp ? EV_PZ : EV_PX)][i] / n_safe);
        } else if (which <= 12) {                           // Lorentz sources
            const int s = (which - 7) / 3, c = (which - 7) % 3;
            T f[3];
            em_lorentz(S, s, x, z, p, f);
            r = (double)f[c];
        } else {
            r = energy_cell_em(S, p.nsp, i);
        }
        out[i] = r;
    }
}

template <typename T>
void launch_derived_em(const State<T>& S, int which, const KParams& p, double* out, hipStream_t st) {
    size_t N = (size_t)p.nx * p.nz;
    int blocks = (int)std::min<size_t>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_derived_em<T>, dim3(blocks), dim3(256), 0, st, S, which, p, out);
}

// --------------------------------------------------------------------------
// Circular mode amplitude (ymgpu2d/analysis/remote_timeseries.py:62-77):
// F(x) = sum_z (Az2 + i Az3) e^{-2 pi i k z / nz}, amp = mean_x |F| / nz, and
// the same for the conjugate signal (Az2 - i Az3).  FP64 accumulation; the
// mean over x is a compensated sum.  One thread per column.
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_mode_amp(const T* a2, const T* a3, int kz, KParams p,
                                                          double* partials) {
    const int nx = p.nx, nz = p.nz;
    const int x = blockIdx.x * kRedThreads + threadIdx.x;
    CSum acc[2] = {CSum{0.0, 0.0}, CSum{0.0, 0.0}};
    if (x < nx) {
        double fr = 0.0, fi = 0.0, gr = 0.0, gi = 0.0;
        for (int z = 0; z < nz; ++z) {
            const size_t i = (size_t)x + (size_t)z * nx;
            const double u = (double)a2[i], v = (double)a3[i];
            double sn, cs;
            sincospi(2.0 * (double)(((long)kz * z) % nz) / (double)nz, &sn, &cs);
            fr += u * cs + v * sn;          // (u + i v)(cos - i sin)
            fi += v * cs - u * sn;
            gr += u * cs - v * sn;          // (u - i v)(cos - i sin)
            gi += -v * cs - u * sn;
        }
        cadd(acc[0], sqrt(fr * fr + fi * fi));
        cadd(acc[1], sqrt(gr * gr + gi * gi));
    }
    block_reduce_store<2>(acc, partials);
}

template <typename T>
void launch_mode_amp(const T* a2, const T* a3, int kz, const KParams& p, double* partials, double* out,
                     hipStream_t st) {
    const int blocks = (p.nx + kRedThreads - 1) / kRedThreads;
    hipLaunchKernelGGL(k_mode_amp<T>, dim3(blocks), dim3(kRedThreads), 0, st, a2, a3, kz, p, partials);
    hipLaunchKernelGGL(k_reduce_final<2>, dim3(1), dim3(kRedThreads), 0, st, partials, blocks, out);
}

// --------------------------------------------------------------------------
// 2-D mode spectrum, FP64 form (plain FP64 sums; the FP32 form runs on MFMA
// in khs_filter_mfma.hip).  Tables have 64 columns: r < K cos, K <= r < 2K sin.
// --------------------------------------------------------------------------
__global__ void k_spec_z_f64(const double* __restrict__ f, const double* __restrict__ tab, double* __restrict__ C,
                             int nx, int nz, int rows) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (x >= nx || r >= rows) return;
    double s = 0.0;
    for (int z = 0; z < nz; ++z) s += f[(size_t)z * nx + x] * tab[(size_t)z * 64 + r];
    C[(size_t)r * nx + x] = s;
}

__global__ void k_spec_x_f64(const double* __restrict__ C, const double* __restrict__ utab, double* __restrict__ D,
                             int nx, int rows) {
    const int s = threadIdx.x, r = blockIdx.x;
    if (r >= rows || s >= rows) return;
    double acc = 0.0;
    for (int x = 0; x < nx; ++x) acc += C[(size_t)r * nx + x] * utab[(size_t)x * 64 + s];
    D[(size_t)r * 64 + s] = acc;
}

void launch_spectrum_f64(const double* f, const double* ztab, const double* xtab, double* C, double* D, int nx,
                         int nz, int rows, hipStream_t st) {
    hipLaunchKernelGGL(k_spec_z_f64, dim3((nx + 127) / 128, rows), dim3(128), 0, st, f, ztab, C, nx, nz, rows);
    hipLaunchKernelGGL(k_spec_x_f64, dim3(rows), dim3(64), 0, st, C, xtab, D, nx, rows);
}

// explicit instantiations
#define KHS_INST(T)                                                                                   \
    template void launch_cell_filters<T>(const ArrayList<T>&, const double*, const int*, const KParams&, hipStream_t, \
                                         int);                                                                     \
    template void launch_zmean_cut<T>(const ArrayList<T>&, int, int, const double*, const int*, const KParams&, hipStream_t); \
    template void launch_column_modes<T, float>(const ArrayList<T>&, const float*, const float*, const float*, int, int, double*, const KParams&, hipStream_t); \
    template void launch_column_modes<T, double>(const ArrayList<T>&, const double*, const double*, const double*, int, int, double*, const KParams&, hipStream_t); \
    template void launch_hyperdiffuse<T>(const ArrayList<T>&, double, const KParams&, hipStream_t); \
    template void launch_enforce_boundaries<T>(const State<T>&, const KParams&, int, hipStream_t);  \
    template void launch_energy<T>(const State<T>&, const KParams&, double*, double*, hipStream_t);  \
    template void launch_gauss<T>(const State<T>&, const KParams&, double*, double*, hipStream_t);   \
    template void launch_derived<T>(const State<T>&, int, const KParams&, double*, hipStream_t);   \
    template void launch_energy_em<T>(const State<T>&, const KParams&, double*, double*, hipStream_t); \
    template void launch_derived_em<T>(const State<T>&, int, const KParams&, double*, hipStream_t); \
    template void launch_mode_amp<T>(const T*, const T*, int, const KParams&, double*, double*, hipStream_t);
KHS_INST(double)
KHS_INST(float)
#undef KHS_INST

}  // namespace khs
