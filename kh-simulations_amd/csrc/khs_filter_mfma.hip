// khs_filter_mfma.hip — the keep-|kz|<=K column filter of the FP32 precision
// form as two FP32 MFMA contractions along z (gfx950 v_mfma_f32_16x16x4_f32,
// FP32 accumulation).
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
// MFMA operand layout (16x16x4 f32, wave64): lane l = i + 16 k holds A[i][k];
// lane l = j + 16 k holds B[k][j]; the result element e of lane l is
// D[4 (l / 16) + e][l % 16].  The 64 columns of a block are assigned to the
// four 16-wide N tiles as column x0 + 4 n + j (tile j, index n), so that one
// 16-byte load or store per lane covers all four tiles.
#include "khs_device.h"
#include "khs_launch.h"

#include <cmath>

namespace khs {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int KF_ROWS = 16;        // mode rows (cos 1..K, sin 1..K, K <= 8)
constexpr int KF_ZR = 256;         // z rows per synthesis block

// Coefficients of one 64-column strip of one array; the 4 waves split z.
__global__ __launch_bounds__(256) void k_keep_coef_mfma(ArrayList<float> arrs, const float* __restrict__ tab,
                                                        float* __restrict__ coef, int nx, int nz) {
    __shared__ float red[4][KF_ROWS][64];
    __shared__ float redm[4][4][64];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int n = l & 15, kq = l >> 4;
    const int x0 = blockIdx.x * 64;
    const int a = blockIdx.y;
    const float* __restrict__ v = arrs.a[a];
    const int col = x0 + 4 * n;
    const bool ok = col < nx;                      // nx % 4 == 0: all four or none
    f4 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = f4{0.f, 0.f, 0.f, 0.f};
    f4 msum = f4{0.f, 0.f, 0.f, 0.f};
    const int zper = nz >> 2;                      // nz % 16 == 0
    const int zb = wave * zper, ze = zb + zper;
    for (int z0 = zb; z0 < ze; z0 += 4) {
        const int z = z0 + kq;
        const f4 b = ok ? *reinterpret_cast<const f4*>(v + (size_t)z * nx + col) : f4{0.f, 0.f, 0.f, 0.f};
        const float at = tab[(size_t)z * KF_ROWS + n];       // A[i = n][k = kq]
        acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(at, b.x, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(at, b.y, acc[1], 0, 0, 0);
        acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(at, b.z, acc[2], 0, 0, 0);
        acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(at, b.w, acc[3], 0, 0, 0);
        msum += b;
    }
    // lane holds, for tile j, rows 4 kq + e of column 4 n + j
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) red[wave][4 * kq + e][4 * n + j] = acc[j][e];
#pragma unroll
    for (int j = 0; j < 4; ++j) redm[wave][kq][4 * n + j] = msum[j];
    __syncthreads();
    const float w = 2.0f / (float)nz, w0 = 1.0f / (float)nz;
    const size_t base = (size_t)a * (KF_ROWS + 1);
    for (int o = threadIdx.x; o < KF_ROWS * 64; o += 256) {
        const int r = o >> 6, c = o & 63;
        if (x0 + c >= nx) continue;
        const float s = (red[0][r][c] + red[1][r][c]) + (red[2][r][c] + red[3][r][c]);
        coef[(base + 1 + r) * nx + x0 + c] = s * w;
    }
    if (threadIdx.x < 64) {
        const int c = threadIdx.x;
        if (x0 + c < nx) {
            float s = 0.f;
            for (int q = 0; q < 16; ++q) s += redm[q >> 2][q & 3][c];
            coef[base * nx + x0 + c] = s * w0;
        }
    }
}

// Synthesis of KF_ZR rows of one 64-column strip; each wave does 16-row tiles.
__global__ __launch_bounds__(256) void k_keep_synth_mfma(ArrayList<float> arrs, const float* __restrict__ tab,
                                                         const float* __restrict__ coef, int nx, int nz) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int n = l & 15, kq = l >> 4;
    const int x0 = blockIdx.x * 64;
    const int a = blockIdx.y;
    const int zr0 = blockIdx.z * KF_ZR;
    float* __restrict__ v = arrs.a[a];
    const int col = x0 + 4 * n;
    // Lanes past the last column still take part in the MFMAs (their A values
    // feed every column); they contribute B = 0 and store nothing.
    const bool ok = col < nx;
    const size_t base = (size_t)a * (KF_ROWS + 1);
    // B[k = kq][n] of k-step s is C[4 kq + s][column 4 n + j]: 16 values per lane
    const f4 zero = f4{0.f, 0.f, 0.f, 0.f};
    f4 bc[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
        bc[s] = ok ? *reinterpret_cast<const f4*>(coef + (base + 1 + 4 * kq + s) * nx + col) : zero;
    const f4 mean = ok ? *reinterpret_cast<const f4*>(coef + base * nx + col) : zero;
    const int zend = min(zr0 + KF_ZR, nz);
    for (int zt = zr0 + 16 * wave; zt < zend; zt += 64) {
        // A[i = n][k = kq] of k-step s is T[zt + n][4 kq + s]
        const f4 at = *reinterpret_cast<const f4*>(tab + (size_t)(zt + n) * KF_ROWS + 4 * kq);
        f4 acc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = f4{mean[j], mean[j], mean[j], mean[j]};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const float as = at[s];
            acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(as, bc[s].x, acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(as, bc[s].y, acc[1], 0, 0, 0);
            acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(as, bc[s].z, acc[2], 0, 0, 0);
            acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(as, bc[s].w, acc[3], 0, 0, 0);
        }
        if (ok) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const f4 o = f4{acc[0][e], acc[1][e], acc[2][e], acc[3][e]};
                *reinterpret_cast<f4*>(v + (size_t)(zt + 4 * kq + e) * nx + col) = o;
            }
        }
    }
}

bool keep_mfma_supported(int nx, int nz, int K) {
    return K >= 1 && 2 * K <= KF_ROWS && nx % 4 == 0 && nz % 16 == 0 && nz >= 16;
}

void fill_keep_mfma_table(int nz, int K, float* tab) {
    for (int z = 0; z < nz; ++z)
        for (int r = 0; r < KF_ROWS; ++r) {
            double v = 0.0;
            if (r < 2 * K) {
                const int m = (r < K) ? r + 1 : r - K + 1;
                const double ph = 2.0 * M_PI * (double)((long)m * z % nz) / (double)nz;
                v = (r < K) ? std::cos(ph) : std::sin(ph);
            }
            tab[(size_t)z * KF_ROWS + r] = (float)v;
        }
}

void launch_keep_mfma(const ArrayList<float>& arrs, const float* tab, float* coef, const KParams& p,
                      hipStream_t st) {
    const int nxb = (p.nx + 63) / 64;
    hipLaunchKernelGGL(k_keep_coef_mfma, dim3(nxb, arrs.n), dim3(256), 0, st, arrs, tab, coef, p.nx, p.nz);
    hipLaunchKernelGGL(k_keep_synth_mfma, dim3(nxb, arrs.n, (p.nz + KF_ZR - 1) / KF_ZR), dim3(256), 0, st, arrs,
                       tab, coef, p.nx, p.nz);
}

}  // namespace khs
