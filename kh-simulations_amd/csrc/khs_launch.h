// khs_launch.h — host-callable launchers of the gfx950 kernels.
#pragma once

#include "khs_device.h"

namespace khs {

// step kernels (khs_step_kernels.hip)
// Tile variants (bit-identical results): x segment 64/128/256 cells
// (variant % 3) with 1/4/16 rows per workgroup ((variant / 3) % 3), all
// variables or one at a time through LDS (variant / 9); z chunk 16/32/64 rows
// (variant % 3) without / with next-row prefetch ((variant / 3) % 2), one lane
// per column or two waves per column (variable groups, variant / 6).  x
// variants 18 / 19: register form, 4 / 8 cells per lane, lane neighbours by
// DPP (YM model).
constexpr int kNumXVariants = 20;
constexpr int kNumZVariants = 15;          // 12..14: 16/32/64-row chunks, loads two rows ahead
inline int z_prefetch_of(int v) { return v >= 12 ? 2 : (v / 3) % 2; }
inline int z_groups_of(int v) { return v >= 12 ? 0 : v / 6; }
constexpr int kNumMaxwellVariants = 11;
// M: YMModel (ymgpu2d two-beam SU(2) fluid) or EMModel (emhdgpu2d Abelian fluid)
template <typename T, typename M>
void launch_fct_x(const State<T>& S, const Stage<T>& out, const KParams& p, int sp0, int nsp,
                  int variant, hipStream_t st);
template <typename T, typename M>
void launch_fct_z(const State<T>& S, const Stage<T>& X, const State<T>& O, const KParams& p,
                  int sp0, int nsp, int variant, hipStream_t st);
template <typename T>
void launch_emhd_maxwell(const State<T>& S, const State<T>& O, const KParams& p, int variant, hipStream_t st);
constexpr int kNumEmMaxwellVariants = 4;   // 0: 64 x 16 tiles; 1/2/3: z march, 4/8/16-row chunks
// EMHD energy (main_coupled.cu:197-222): out[0] = sum over cells
template <typename T>
void launch_energy_em(const State<T>& S, const KParams& p, double* partials, double* out, hipStream_t st);
// EMHD derived arrays: 0..2 J, 3..6 (ux, uz) per fluid, 7..12 (sx, sy, sz) per fluid, 13 energy density
template <typename T>
void launch_derived_em(const State<T>& S, int which, const KParams& p, double* out, hipStream_t st);
// fused x + z sweep of both species: state S -> fluid of O (variants: see
// launch_fct_xz in khs_step_kernels.hip)
constexpr int kNumXZVariants = 15;
// variants from this one on take no warm terms (warm_T = 0 only)
constexpr int kFirstColdXZVariant = 13;
// EMHD fused x+z sweep of species 0..nsp-1: variants 0/1/2/3 (16/32/64/8-row chunks), 4/5/6
// (two wavefronts per species, 8/16/32-row chunks)
constexpr int kNumXZEmVariants = 7;
template <typename T>
void launch_fct_xz_em(const State<T>& S, const State<T>& O, const KParams& p, int nsp, int variant, hipStream_t st);
template <typename T>
void launch_fct_xz(const State<T>& S, const State<T>& O, const KParams& p, int variant, hipStream_t st);
// decay/action: per-column sponge / xi_cut (nullptr when neither is active);
// variant 0/2/3: 64x16 / 64x4 / 64x8 tiles sharing the new E through LDS,
// 1: one thread per cell recomputing its neighbours' E, 4..10: z march, one
// lane per column, chunks of 16/32/16/8/8/4/2 rows with / with / without /
// with / without / with / with next-row prefetch; all bit-identical
// fused z sweep (both species) + Maxwell of the YM model: stage X and state S
// -> fluid and fields of O; variant 1 / 2 / 3 = 16 / 32 / 64-row chunks
constexpr int kNumZMVariants = 4;          // 0 = split z sweep and Maxwell kernels
template <typename T>
void launch_fct_zmx(const State<T>& S, const Stage<T>& X, const State<T>& O, const KParams& p, const double* decay,
                    const int* action, int variant, hipStream_t st);
template <typename T>
void launch_maxwell(const State<T>& S, const State<T>& O, const KParams& p, const double* decay,
                    const int* action, int variant, hipStream_t st);
// Maxwell (z-march variants 4..10) that also writes one compensated partial
// sum of the new state's energy per wavefront into epart; returns their
// number, 0 (nothing launched) for the other variants.
template <typename T>
int launch_maxwell_energy(const State<T>& S, const State<T>& O, const KParams& p, const double* decay,
                          const int* action, int variant, double* epart, hipStream_t st);
// Maxwell of step n (with the sponge / xi_cut of its colour-2/3 fields) and
// the fused x + z sweeps of step n + 1 (with the same filter applied to the
// charges it reads) in one launch, YM model without column filters.  I:
// fields of step n, fluid after the sweeps of step n; O: fields of step
// n + 1, fluid after the sweeps of step n + 1.  Variants 0/1/2 = 16/32/64-row
// chunks with two Maxwell wavefronts per workgroup, 3/4/5 with four, 6/7 with
// four and chunks sized so that the tiles fill the device's resident
// workgroups once / twice, 8 with two and one round.
constexpr int kNumMxzVariants = 9;
template <typename T>
void launch_maxwell_sweep(const State<T>& I, const State<T>& O, const KParams& p, const double* decay,
                          const int* action, int variant, hipStream_t st);

// filters (khs_aux_kernels.hip)
constexpr int kMaxFilterArrays = 24;
template <typename T>
struct ArrayList {
    T* a[kMaxFilterArrays];
    int n;
};
// Per-column action of the sponge / xi_cut kernels: 0 none, 1 multiply by
// decay[x], 2 zero (action may be null: no action); the last nzero arrays of
// the list are set to zero everywhere.
template <typename T>
void launch_cell_filters(const ArrayList<T>& arrs, const double* decay, const int* action,
                         const KParams& p, hipStream_t st, int nzero = 0);
// z-mean removal of arrays [0, nmean) (the last nact after the per-column
// action of launch_cell_filters, when action is given) and zeroing of arrays
// [nmean, n), one workgroup per 64-column strip and array.
template <typename T>
void launch_zmean_cut(const ArrayList<T>& arrs, int nmean, int nact, const double* decay, const int* action,
                      const KParams& p, hipStream_t st);
// Column DFT filter.  coef scratch: arrays * 2 * nmodes * nx doubles.
//   project == 0: v -= sum_m (a_m c_m(z) + b_m s_m(z))   (reference band subtraction)
//   project == 1: v  = sum_m (a_m c_m(z) + b_m s_m(z))   (keep-modes projection)
// a_m = weight[m] * sum_z v c_m(z), b_m likewise with s_m; tables are [m][nz].
template <typename T, typename TW>
void launch_column_modes(const ArrayList<T>& arrs, const TW* ctab, const TW* stab,
                         const TW* weight, int nmodes, int project, double* coef,
                         const KParams& p, hipStream_t st);
// keep-|kz|<=K projection on MFMA (khs_filter_mfma.hip).  tab: nz x 16 values
// from fill_keep_mfma_table; coef scratch: arrays * 17 * nx values.
bool keep_mfma_supported(int nx, int nz, int K);
template <typename T>
void fill_keep_mfma_table(int nz, int K, T* tab);
template <typename T>
void launch_keep_mfma(const ArrayList<T>& arrs, const T* tab, T* coef, const KParams& p, hipStream_t st,
                      int nomean = 0);
// 2-D mode spectrum (kx, kz = 1..K): tables of 64 columns per grid point
// (cos for r < K, sin for K <= r < 2K), C scratch 64 x nx, D output 64 x 64.
bool spectrum_mfma_supported(int nx, int nz, int K);
void fill_spectrum_table(int n, int K, float* tab);
// scratch of the FP32 spectrum: C holds spectrum_z_slices(nx, nz) x 64 x nx
// partial z contractions, P spectrum_x_blocks(nx) x 64 x 64 partial x
// contractions
int spectrum_x_blocks(int nx);
int spectrum_z_slices(int nx, int nz);
void launch_spectrum_mfma(const float* f, const float* ztab, const float* xtab, float* C, float* P, float* D, int nx,
                          int nz, hipStream_t st);
void launch_spectrum_f64(const double* f, const double* ztab, const double* xtab, double* C, double* D, int nx,
                         int nz, int rows, hipStream_t st);
template <typename T>
void launch_hyperdiffuse(const ArrayList<T>& arrs, double mu, const KParams& p, hipStream_t st);
template <typename T>
void launch_enforce_boundaries(const State<T>& S, const KParams& p, int emhd, hipStream_t st);

// diagnostics
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;
// The compensated sum of n (sum, compensation) pairs into out[0]; scratch
// holds ceil(n / (4 kRedThreads)) pairs for long lists.
void launch_reduce_partials(const double* partials, int n, double* scratch, double* out, hipStream_t st);
// Energy: out[0] = sum, out[1] = nonfinite flag (>0 if NaN/Inf seen).
template <typename T>
void launch_energy(const State<T>& S, const KParams& p, double* partials, double* out,
                   hipStream_t st);
// Gauss residual sums: out[0..5] = sum G^2 (two signs), out[6] = sum (sum|E|)^2.
template <typename T>
void launch_gauss(const State<T>& S, const KParams& p, double* partials, double* out,
                  hipStream_t st);
// Circular amplitude of (a2 + i a3) at z mode kz: out[0] = sum_x |F|, out[1] =
// the same for the conjugate signal (the caller divides by nx * nz).
template <typename T>
void launch_mode_amp(const T* a2, const T* a3, int kz, const KParams& p, double* partials, double* out,
                     hipStream_t st);
// Derived per-cell array (ids of khs.h >= 32) evaluated from S into out (double).
template <typename T>
void launch_derived(const State<T>& S, int id, const KParams& p, double* out, hipStream_t st);

// --------------------------------------------------------------------------
// Packed runs: many independent runs of one batch advanced by shared launches
// (one fused-sweep launch, one Maxwell launch and one filter launch per step
// for all of them).  A launch's tiles are the runs' tiles one run after
// another; a workgroup finds its run in `begin` (prefix offsets, n + 1
// entries, device memory) and steps it with the run's own state pointers and
// parameters, exactly as the run's own launch would.
// --------------------------------------------------------------------------
template <typename T>
struct PackedRun {
    State<T> S, O;                 // the step's input and output state sets
    KParams p;
    const double* decay;           // xi_cut / sponge column action, or null
    const int* action;
    int ngx;                       // fused sweep: 58-column groups (2 ngx nch tiles)
    int nwx;                       // Maxwell march: 63-column groups (nwx nchm waves)
    ArrayList<T> filt;             // filter stage: arrays of the run
    int fmode;                     // 0 none, 1 charge cut / sponge on filt, 2 z-mean + cut + colour-1 zero
    int nmean, nact;               // fmode 2: see launch_zmean_cut
    int nstrip;                    // filter work items per array: 64-column strips
};
// Packed launch shapes (chunk rows of the fused sweep, of the Maxwell march;
// two sweep tiles per workgroup unless noted): 0: 16, 8; 1: 32, 4; 2: 32, 8;
// 3: 16, 4; 4: 32, 2; 5: 16, 2; 6: 32, 4 and 7: 8, 4 with one sweep tile
// per workgroup.
constexpr int kPackedShapes = 8;
inline int packed_xz_chunk(int shape) {
    static const int rows[kPackedShapes] = {16, 32, 32, 16, 32, 16, 32, 8};
    return rows[shape];
}
inline int packed_mx_chunk(int shape) {
    static const int rows[kPackedShapes] = {8, 4, 8, 4, 2, 2, 4, 4};
    return rows[shape];
}
// begin: device arrays of n + 1 tile offsets per launch kind (fused-sweep
// tiles, Maxwell waves, filter work items) for the shape
template <typename T>
void launch_filters_packed(const PackedRun<T>* runs, const int* begin, int n, int total, hipStream_t st);
template <typename T>
void launch_packed_step(const PackedRun<T>* runs, const int* xz_begin, const int* mx_begin, const int* f_begin,
                        int n, int xz_total, int mx_total, int f_total, int shape, hipStream_t st);

}  // namespace khs
