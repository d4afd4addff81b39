// khs_filter_mfma.hip — the keep-|kz|<=K column filter as two MFMA
// contractions along z: FP32 form on v_mfma_f32_16x16x4_f32 (FP32
// accumulation), FP64 form on v_mfma_f64_16x16x4_f64.
//
// For every filtered array V (nz rows x nx columns, x contiguous) and the
// table T[z][r] (r < K: cos((r+1) th_z), K <= r < 2K: sin((r-K+1) th_z),
// th_z = 2 pi z / nz, zero rows above 2K):
//
//   coefficients  C[r][x] = (2/nz) sum_z T[z][r] V[z][x],  mean[x] = (1/nz) sum_z V[z][x]
//   synthesis     V[z][x] = mean[x] + sum_r T[z][r] C[r][x]
//
// which is the keep-modes projection of khs_aux_kernels.hip (k_col_coef /
// k_col_apply with project = 1) for K <= 8.  Both passes stream V once, so the
// filter is HBM-bound (R + W of each array); the MFMA work is 2 x 16 x 4 flops
// per element and pass.
//
// MFMA operand layout (16x16x4, wave64): lane l = i + 16 k holds A[i][k];
// lane l = j + 16 k holds B[k][j]; result element e of lane l is
// D[4 (l / 16) + e][l % 16] for f32 and D[l / 16 + 4 e][l % 16] for f64
// (both measured on gfx950).  The 64 columns of a block are assigned to the
// four 16-wide N tiles as column x0 + 4 n + j (tile j, index n), so that one
// 16-byte load or store per lane covers all four tiles.
#include "khs_device.h"
#include "khs_launch.h"

#include <cmath>

namespace khs {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

template <typename T> struct V4;
template <> struct V4<float> { typedef f4 type; };
template <> struct V4<double> { typedef d4 type; };

// Row of D held in result element e by the lanes of group kq = l / 16.
__device__ __forceinline__ int drow(float, int kq, int e) { return 4 * kq + e; }
__device__ __forceinline__ int drow(double, int kq, int e) { return kq + 4 * e; }

__device__ __forceinline__ f4 mfma16(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ d4 mfma16(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int KF_ROWS = 16;        // mode rows (cos 1..K, sin 1..K, K <= 8)
constexpr int KF_ZR = 256;         // z rows per synthesis block

// Coefficients of one 64-column strip of one array; the 4 waves split z.
template <typename T>
__global__ __launch_bounds__(256) void k_keep_coef_mfma(ArrayList<T> arrs, const T* __restrict__ tab,
                                                        T* __restrict__ coef, int nx, int nz) {
    typedef typename V4<T>::type v4;
    __shared__ T red[4][KF_ROWS][64];
    __shared__ T redm[4][4][64];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int n = l & 15, kq = l >> 4;
    const int x0 = blockIdx.x * 64;
    const int a = blockIdx.y;
    const T* __restrict__ v = arrs.a[a];
    const int col = x0 + 4 * n;
    const bool ok = col < nx;                      // nx % 4 == 0: all four or none
    const v4 zero = v4{T(0), T(0), T(0), T(0)};
    v4 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = zero;
    v4 msum = zero;
    const int zper = nz >> 2;                      // nz % 16 == 0
    const int zb = wave * zper, ze = zb + zper;
    for (int z0 = zb; z0 < ze; z0 += 4) {
        const int z = z0 + kq;
        const v4 b = ok ? *reinterpret_cast<const v4*>(v + (size_t)z * nx + col) : zero;
        const T at = tab[(size_t)z * KF_ROWS + n];           // A[i = n][k = kq]
        acc[0] = mfma16(at, b.x, acc[0]);
        acc[1] = mfma16(at, b.y, acc[1]);
        acc[2] = mfma16(at, b.z, acc[2]);
        acc[3] = mfma16(at, b.w, acc[3]);
        msum += b;
    }
    // lane holds, for tile j, rows drow(kq, e) of column 4 n + j
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) red[wave][drow(T(0), kq, e)][4 * n + j] = acc[j][e];
#pragma unroll
    for (int j = 0; j < 4; ++j) redm[wave][kq][4 * n + j] = msum[j];
    __syncthreads();
    const T w = T(2) / T(nz), w0 = T(1) / T(nz);
    const size_t base = (size_t)a * (KF_ROWS + 1);
    for (int o = threadIdx.x; o < KF_ROWS * 64; o += 256) {
        const int r = o >> 6, c = o & 63;
        if (x0 + c >= nx) continue;
        const T s = (red[0][r][c] + red[1][r][c]) + (red[2][r][c] + red[3][r][c]);
        coef[(base + 1 + r) * nx + x0 + c] = s * w;
    }
    if (threadIdx.x < 64) {
        const int c = threadIdx.x;
        if (x0 + c < nx) {
            T s = T(0);
            for (int q = 0; q < 16; ++q) s += redm[q >> 2][q & 3][c];
            coef[base * nx + x0 + c] = s * w0;
        }
    }
}

// Synthesis of KF_ZR rows of one 64-column strip; each wave does 16-row tiles.
template <typename T>
__global__ __launch_bounds__(256) void k_keep_synth_mfma(ArrayList<T> arrs, const T* __restrict__ tab,
                                                         const T* __restrict__ coef, int nx, int nz) {
    typedef typename V4<T>::type v4;
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int n = l & 15, kq = l >> 4;
    const int x0 = blockIdx.x * 64;
    const int a = blockIdx.y;
    const int zr0 = blockIdx.z * KF_ZR;
    T* __restrict__ v = arrs.a[a];
    const int col = x0 + 4 * n;
    // Lanes past the last column still take part in the MFMAs (their A values
    // feed every column); they contribute B = 0 and store nothing.
    const bool ok = col < nx;
    const size_t base = (size_t)a * (KF_ROWS + 1);
    // B[k = kq][n] of k-step s is C[4 kq + s][column 4 n + j]: 16 values per lane
    const v4 zero = v4{T(0), T(0), T(0), T(0)};
    v4 bc[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
        bc[s] = ok ? *reinterpret_cast<const v4*>(coef + (base + 1 + 4 * kq + s) * nx + col) : zero;
    const v4 mean = ok ? *reinterpret_cast<const v4*>(coef + base * nx + col) : zero;
    const int zend = min(zr0 + KF_ZR, nz);
    for (int zt = zr0 + 16 * wave; zt < zend; zt += 64) {
        // A[i = n][k = kq] of k-step s is T[zt + n][4 kq + s]
        const v4 at = *reinterpret_cast<const v4*>(tab + (size_t)(zt + n) * KF_ROWS + 4 * kq);
        v4 acc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = v4{mean[j], mean[j], mean[j], mean[j]};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const T as = at[s];
            acc[0] = mfma16(as, bc[s].x, acc[0]);
            acc[1] = mfma16(as, bc[s].y, acc[1]);
            acc[2] = mfma16(as, bc[s].z, acc[2]);
            acc[3] = mfma16(as, bc[s].w, acc[3]);
        }
        if (ok) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const v4 o = v4{acc[0][e], acc[1][e], acc[2][e], acc[3][e]};
                *reinterpret_cast<v4*>(v + (size_t)(zt + drow(T(0), kq, e)) * nx + col) = o;
            }
        }
    }
}

bool keep_mfma_supported(int nx, int nz, int K) {
    return K >= 1 && 2 * K <= KF_ROWS && nx % 4 == 0 && nz % 16 == 0 && nz >= 16;
}

template <typename T>
void fill_keep_mfma_table(int nz, int K, T* tab) {
    for (int z = 0; z < nz; ++z)
        for (int r = 0; r < KF_ROWS; ++r) {
            double v = 0.0;
            if (r < 2 * K) {
                const int m = (r < K) ? r + 1 : r - K + 1;
                const double ph = 2.0 * M_PI * (double)((long)m * z % nz) / (double)nz;
                v = (r < K) ? std::cos(ph) : std::sin(ph);
            }
            tab[(size_t)z * KF_ROWS + r] = (T)v;
        }
}

template <typename T>
void launch_keep_mfma(const ArrayList<T>& arrs, const T* tab, T* coef, const KParams& p, hipStream_t st) {
    const int nxb = (p.nx + 63) / 64;
    hipLaunchKernelGGL(k_keep_coef_mfma<T>, dim3(nxb, arrs.n), dim3(256), 0, st, arrs, tab, coef, p.nx, p.nz);
    hipLaunchKernelGGL(k_keep_synth_mfma<T>, dim3(nxb, arrs.n, (p.nz + KF_ZR - 1) / KF_ZR), dim3(256), 0, st,
                       arrs, tab, coef, p.nx, p.nz);
}

template void fill_keep_mfma_table<float>(int, int, float*);
template void fill_keep_mfma_table<double>(int, int, double*);
template void launch_keep_mfma<float>(const ArrayList<float>&, const float*, float*, const KParams&, hipStream_t);
template void launch_keep_mfma<double>(const ArrayList<double>&, const double*, double*, const KParams&,
                                       hipStream_t);

}  // namespace khs
