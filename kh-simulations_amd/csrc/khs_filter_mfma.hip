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
// 16-byte load or store per lane covers all four tiles; the FP64 synthesis
// uses column x0 + 16 j + n instead (whole 128-byte lines per 8-byte store).
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

// Coefficients of one strip of 64 * Q columns of one array; the 4 waves split
// z.  Lane n of a 16-lane row group holds columns x0 + 4 (Q n + q) + j
// (q < Q, j < 4) with Q 16-byte loads, tile 4 q + j.
template <typename T, int Q>
__global__ __launch_bounds__(256) void k_keep_coef_mfma(ArrayList<T> arrs, const T* __restrict__ tab,
                                                        T* __restrict__ coef, int nx, int nz) {
    typedef typename V4<T>::type v4;
    constexpr int W = 64 * Q;                      // columns per strip
    __shared__ T red[4][KF_ROWS][W];
    __shared__ T redm[4][4][W];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int n = l & 15, kq = l >> 4;
    const int x0 = blockIdx.x * W;
    const int a = blockIdx.y;
    const T* __restrict__ v = arrs.a[a];
    bool ok[Q];                                    // nx % 4 == 0: all four or none
#pragma unroll
    for (int q = 0; q < Q; ++q) ok[q] = x0 + 4 * (Q * n + q) < nx;
    const v4 zero = v4{T(0), T(0), T(0), T(0)};
    v4 acc[4 * Q];
#pragma unroll
    for (int j = 0; j < 4 * Q; ++j) acc[j] = zero;
    v4 msum[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) msum[q] = zero;
    const int zper = nz >> 2;                      // nz % 16 == 0
    const int zb = wave * zper, ze = zb + zper;
    auto ld = [&](int z, int q) {
        return ok[q] ? *reinterpret_cast<const v4*>(v + (size_t)z * nx + x0 + 4 * (Q * n + q)) : zero;
    };
    auto step = [&](T at, const v4 (&b)[Q]) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            acc[4 * q + 0] = mfma16(at, b[q].x, acc[4 * q + 0]);
            acc[4 * q + 1] = mfma16(at, b[q].y, acc[4 * q + 1]);
            acc[4 * q + 2] = mfma16(at, b[q].z, acc[4 * q + 2]);
            acc[4 * q + 3] = mfma16(at, b[q].w, acc[4 * q + 3]);
            msum[q] += b[q];
        }
    };
    // four 4-row steps per iteration with their loads issued first (more
    // loads in flight per wave); the MFMA and sum order is the one-step order
    int z0 = zb;
    for (; z0 + 16 <= ze; z0 += 16) {
        v4 b[4][Q];
        T at[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int z = z0 + 4 * u + kq;
#pragma unroll
            for (int q = 0; q < Q; ++q) b[u][q] = ld(z, q);
            at[u] = tab[(size_t)z * KF_ROWS + n];              // A[i = n][k = kq]
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) step(at[u], b[u]);
    }
    for (; z0 < ze; z0 += 4) {
        const int z = z0 + kq;
        v4 b[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) b[q] = ld(z, q);
        step(tab[(size_t)z * KF_ROWS + n], b);
    }
    // lane holds, for tile 4 q + j, rows drow(kq, e) of column 4 (Q n + q) + j
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int e = 0; e < 4; ++e) red[wave][drow(T(0), kq, e)][4 * (Q * n + q) + j] = acc[4 * q + j][e];
            redm[wave][kq][4 * (Q * n + q) + j] = msum[q][j];
        }
    __syncthreads();
    const T w = T(2) / T(nz), w0 = T(1) / T(nz);
    const size_t base = (size_t)a * (KF_ROWS + 1);
    for (int o = threadIdx.x; o < KF_ROWS * W; o += 256) {
        const int r = o / W, c = o % W;
        if (x0 + c >= nx) continue;
        const T s = (red[0][r][c] + red[1][r][c]) + (red[2][r][c] + red[3][r][c]);
        coef[(base + 1 + r) * nx + x0 + c] = s * w;
    }
    for (int c = threadIdx.x; c < W; c += 256) {
        if (x0 + c < nx) {
            T s = T(0);
            for (int qq = 0; qq < 16; ++qq) s += redm[qq >> 2][qq & 3][c];
            coef[base * nx + x0 + c] = s * w0;
        }
    }
}

// Synthesis of KF_ZR rows of one 64-column strip; each wave does 16-row tiles.
// Arrays a < nomean are synthesised without their mean (kz = 0 removed as
// well: the z-mean removal of suppress_kz0 folded into the projection).
template <typename T>
__global__ __launch_bounds__(256) void k_keep_synth_mfma(ArrayList<T> arrs, const T* __restrict__ tab,
                                                         const T* __restrict__ coef, int nx, int nz, int nomean) {
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
    const bool with_mean = a >= nomean;
    const size_t base = (size_t)a * (KF_ROWS + 1);
    if constexpr (sizeof(T) == 8) {
        // FP64: tile j holds columns x0 + 16 j + n, so that one 8-byte store
        // per lane writes whole 128-byte lines, non-temporal (with the FP32
        // map a 32-byte lane value splits into two stores that each cover
        // half of every 32 bytes, twice as slow non-temporal); the sums of
        // a column are those of the FP32 map.  4096x2048: step -2.5%
        T bs[4][4], mj[4];
        bool okj[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int cj = x0 + 16 * j + n;
            okj[j] = cj < nx;
            mj[j] = (okj[j] && with_mean) ? coef[base * nx + cj] : T(0);
#pragma unroll
            for (int s = 0; s < 4; ++s) bs[s][j] = okj[j] ? coef[(base + 1 + 4 * kq + s) * nx + cj] : T(0);
        }
        const int zend = min(zr0 + KF_ZR, nz);
        for (int zt = zr0 + 16 * wave; zt < zend; zt += 64) {
            const v4 at = *reinterpret_cast<const v4*>(tab + (size_t)(zt + n) * KF_ROWS + 4 * kq);
            v4 acc[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[j] = v4{mj[j], mj[j], mj[j], mj[j]};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const T as = at[s];
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[j] = mfma16(as, bs[s][j], acc[j]);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (okj[j])
                        __builtin_nontemporal_store(acc[j][e],
                                                    v + (size_t)(zt + drow(T(0), kq, e)) * nx + x0 + 16 * j + n);
        }
        return;
    }
    // B[k = kq][n] of k-step s is C[4 kq + s][column 4 n + j]: 16 values per lane
    const v4 zero = v4{T(0), T(0), T(0), T(0)};
    v4 bc[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
        bc[s] = ok ? *reinterpret_cast<const v4*>(coef + (base + 1 + 4 * kq + s) * nx + col) : zero;
    const v4 mean = (ok && with_mean) ? *reinterpret_cast<const v4*>(coef + base * nx + col) : zero;
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
                // one 16-byte non-temporal store per lane, 256 contiguous bytes
                // per row group (4096x2048: the next step's sweep and energy
                // pass read faster, step -1.8%)
                __builtin_nontemporal_store(o, reinterpret_cast<v4*>(v + (size_t)(zt + drow(T(0), kq, e)) * nx + col));
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
void launch_keep_mfma(const ArrayList<T>& arrs, const T* tab, T* coef, const KParams& p, hipStream_t st,
                      int nomean) {
    const int nxb = (p.nx + 63) / 64;
    // 128-column strips: a row group reads 512 (FP32) or 1024 (FP64) contiguous
    // bytes per row; 4096x2048, 21 arrays: FP32 pass 7% faster than 64-column
    // strips and than 256-column ones (LDS-bound occupancy), FP64 2%
    constexpr int Q = 2;
    hipLaunchKernelGGL((k_keep_coef_mfma<T, Q>), dim3((p.nx + 64 * Q - 1) / (64 * Q), arrs.n), dim3(256), 0, st,
                       arrs, tab, coef, p.nx, p.nz);
    hipLaunchKernelGGL(k_keep_synth_mfma<T>, dim3(nxb, arrs.n, (p.nz + KF_ZR - 1) / KF_ZR), dim3(256), 0, st,
                       arrs, tab, coef, p.nx, p.nz, nomean);
}

template void fill_keep_mfma_table<float>(int, int, float*);
template void fill_keep_mfma_table<double>(int, int, double*);
template void launch_keep_mfma<float>(const ArrayList<float>&, const float*, float*, const KParams&, hipStream_t,
                                     int);
template void launch_keep_mfma<double>(const ArrayList<double>&, const double*, double*, const KParams&,
                                       hipStream_t, int);

}  // namespace khs

namespace khs {

// --------------------------------------------------------------------------
// 2-D mode spectrum of one array on MFMA (FP32 form):
//   C[r][x] = sum_z T[z][r] f[z][x]            (z contraction, MT x 16 rows)
//   D[r][s] = sum_x C[r][x] U[x][s]            (x contraction, 64 columns)
// with T[z][r] = cos/sin(2 pi kz z / nz) for kz = 1..K (rows r < K: cos,
// K <= r < 2K: sin) and U[x][s] likewise along x.  D holds the four real
// products (cos-cos, cos-sin, sin-cos, sin-sin) from which the complex
// coefficients sum f e^{-2 pi i (kx x / nx + kz z / nz)} follow.
// --------------------------------------------------------------------------
constexpr int SP_ROWS = 64;      // 2K <= 64 (K <= 32)

// z contraction: one 64-column strip and one of gridDim.y z slices per
// block, the 4 waves split the slice; slice y writes its partial rows to
// C + y * 64 * nx (the x contraction adds the slices in order).
__global__ __launch_bounds__(256) void k_spec_z_mfma(const float* __restrict__ f, const float* __restrict__ tab,
                                                     float* __restrict__ C, int nx, int nz) {
    __shared__ float red[SP_ROWS][64];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int n = l & 15, kq = l >> 4;
    const int x0 = blockIdx.x * 64;
    const int col = x0 + 4 * n;
    const bool ok = col < nx;
    const f4 zero = f4{0.f, 0.f, 0.f, 0.f};
    f4 acc[4][4];                                   // [M tile][N tile]
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[t][j] = zero;
    const int zslice = nz / (int)gridDim.y;         // nz % (16 * gridDim.y) == 0
    const int zper = zslice >> 2;
    const int zb = (int)blockIdx.y * zslice + wave * zper, ze = zb + zper;
    for (int z0 = zb; z0 < ze; z0 += 4) {
        const int z = z0 + kq;
        const f4 b = ok ? *reinterpret_cast<const f4*>(f + (size_t)z * nx + col) : zero;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float at = tab[(size_t)z * SP_ROWS + 16 * t + n];
            acc[t][0] = mfma16(at, b.x, acc[t][0]);
            acc[t][1] = mfma16(at, b.y, acc[t][1]);
            acc[t][2] = mfma16(at, b.z, acc[t][2]);
            acc[t][3] = mfma16(at, b.w, acc[t][3]);
        }
    }
    // sum the four z quarters through LDS, one wave at a time
    for (int w = 0; w < 4; ++w) {
        if (wave == w) {
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        float& r = red[16 * t + drow(0.f, kq, e)][4 * n + j];
                        r = (w == 0) ? acc[t][j][e] : r + acc[t][j][e];
                    }
        }
        __syncthreads();
    }
    float* __restrict__ Cs = C + (size_t)blockIdx.y * SP_ROWS * nx;
    for (int o = threadIdx.x; o < SP_ROWS * 64; o += 256) {
        const int r = o >> 6, c = o & 63;
        if (x0 + c < nx) Cs[(size_t)r * nx + x0 + c] = red[r][c];
    }
}

// x contraction: D[64][64] = C[64][nx] U[nx][64], one block; wave w owns
// output row tile w and all four column tiles.
// x contraction, split over x: block b sums the 64 columns 64 b .. 64 b + 63
// into its own 64 x 64 partial product (4 waves x 16 rows r, 4 tiles of 16
// columns s); k_spec_x_sum adds the partials in block order.
__global__ __launch_bounds__(256) void k_spec_x_mfma(const float* __restrict__ C, const float* __restrict__ utab,
                                                     float* __restrict__ P, int nx) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int n = l & 15, kq = l >> 4;
    const f4 zero = f4{0.f, 0.f, 0.f, 0.f};
    f4 acc[4] = {zero, zero, zero, zero};
    const int xb = blockIdx.x * 64;
#pragma unroll 4
    for (int x0 = xb; x0 < xb + 64; x0 += 4) {
        const int x = x0 + kq;
        const bool ok = x < nx;
        const float a = ok ? C[(size_t)(16 * wave + n) * nx + x] : 0.f;   // A[i = n][k = kq]
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float b = ok ? utab[(size_t)x * SP_ROWS + 16 * j + n] : 0.f;   // B[k = kq][j' = n]
            acc[j] = mfma16(a, b, acc[j]);
        }
    }
    float* __restrict__ out = P + (size_t)blockIdx.x * SP_ROWS * SP_ROWS;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) out[(size_t)(16 * wave + drow(0.f, kq, e)) * SP_ROWS + 16 * j + n] = acc[j][e];
}

// out[o] = sum over k < nparts of in[k * stride + o], in order of k (one
// output per thread): the z slices of C into slice 0, the x partials into D
__global__ __launch_bounds__(256) void k_spec_sum(const float* in, float* out, int n,
                                                  size_t stride, int nparts) {
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= n) return;
    float s = 0.f;
    for (int k = 0; k < nparts; ++k) s += in[(size_t)k * stride + o];
    out[o] = s;
}

bool spectrum_mfma_supported(int nx, int nz, int K) {
    return K >= 1 && 2 * K <= SP_ROWS && nx % 4 == 0 && nz % 16 == 0;
}

void fill_spectrum_table(int n, int K, float* tab) {
    for (int z = 0; z < n; ++z)
        for (int r = 0; r < SP_ROWS; ++r) {
            double v = 0.0;
            if (r < 2 * K) {
                const int m = (r < K) ? r + 1 : r - K + 1;
                const double ph = 2.0 * M_PI * (double)((long)m * z % n) / (double)n;
                v = (r < K) ? std::cos(ph) : std::sin(ph);
            }
            tab[(size_t)z * SP_ROWS + r] = (float)v;
        }
}

int spectrum_x_blocks(int nx) { return (nx + 63) / 64; }

// z slices of the z contraction: about 256 blocks in all, each wave at
// least 4 rows (nz % (16 * slices) == 0)
int spectrum_z_slices(int nx, int nz) {
    int zs = 1;
    while (zs < 16 && spectrum_x_blocks(nx) * zs < 256 && nz % (16 * 2 * zs) == 0) zs *= 2;
    return zs;
}

void launch_spectrum_mfma(const float* f, const float* ztab, const float* xtab, float* C, float* P, float* D, int nx,
                          int nz, hipStream_t st) {
    const int nb = spectrum_x_blocks(nx), zs = spectrum_z_slices(nx, nz);
    const int nc = SP_ROWS * nx, nd = SP_ROWS * SP_ROWS;
    hipLaunchKernelGGL(k_spec_z_mfma, dim3(nb, zs), dim3(256), 0, st, f, ztab, C, nx, nz);
    if (zs > 1)   // slices 1.. added onto slice 0 in place: each output reads only its own column of parts
        hipLaunchKernelGGL(k_spec_sum, dim3((nc + 255) / 256), dim3(256), 0, st, C, C, nc, (size_t)nc, zs);
    hipLaunchKernelGGL(k_spec_x_mfma, dim3(nb), dim3(256), 0, st, C, xtab, P, nx);
    hipLaunchKernelGGL(k_spec_sum, dim3((nd + 255) / 256), dim3(256), 0, st, P, D, nd, (size_t)nd, nb);
}

}  // namespace khs
