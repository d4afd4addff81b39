/*
 * khs.h — C ABI of the MI355X (gfx950) Kelvin–Helmholtz time-step library.
 *
 * The library advances the ymgpu2d two-fluid SU(2) Yang–Mills plasma (and the
 * emhdgpu2d Abelian EMHD variant) on one GPU.  Everything behind this header is
 * owned by an opaque context; no device pointer crosses the ABI.  Host arrays
 * use the reference layout: index = x + z*nx, x contiguous (FCT_DeviceMath.cuh:12).
 *
 * Every entry point returns an int status (KHS_OK = 0, negative on error) and
 * never exits the process; the reference's CUDA_CHECK -> exit(1) convention
 * (ymgpu2d/YM_Common.cuh:8-11) and config exit(1) paths (YM_Config.cu:17-20,
 * 120-127) become error codes with a message in khs_last_error().
 * One context must be driven by one host thread at a time; separate contexts
 * are independent.
 */
#ifndef KHS_H
#define KHS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KHS_OK            0
#define KHS_ERR_ARG      -1   /* bad argument (null pointer, size mismatch, unknown id) */
#define KHS_ERR_CONFIG   -2   /* config file missing / unparsable / violates a constraint */
#define KHS_ERR_IO       -3   /* file read or write failed */
#define KHS_ERR_DEVICE   -4   /* HIP runtime error, or no GPU */
#define KHS_ERR_STATE    -5   /* call out of order (e.g. z-sweep before x-sweep) */

#define KHS_FP64 0            /* plain FP64 drop-in form */
#define KHS_FP32 1            /* precision form: FP32 bulk, FP64 reductions */

#define KHS_MODEL_YM   0      /* ymgpu2d: two-beam SU(2) Yang-Mills plasma */
#define KHS_MODEL_EMHD 1      /* emhdgpu2d: Abelian electron-MHD fluid */

#define KHS_SPECIES_A 0       /* beam +z, Q = (+1,0,0) at init */
#define KHS_SPECIES_B 1       /* beam -z, Q = (-1,0,0) at init */

/* State array ids (khs_get_array / khs_set_array).  Names accepted by
 * khs_array_id() are the CSV column names of main_ym.cu:291-292 plus the
 * lower-case px/pz spellings. */
enum khs_array {
    KHS_EX1 = 0, KHS_EX2, KHS_EX3,
    KHS_EZ1, KHS_EZ2, KHS_EZ3,
    KHS_BY1, KHS_BY2, KHS_BY3,
    KHS_AZ1, KHS_AZ2, KHS_AZ3,
    KHS_NA, KHS_PXA, KHS_PZA, KHS_Q1A, KHS_Q2A, KHS_Q3A,
    KHS_NB, KHS_PXB, KHS_PZB, KHS_Q1B, KHS_Q2B, KHS_Q3B,
    KHS_N_STATE,                       /* = 24 evolved arrays */
    /* derived, read-only: evaluated from the current state on request */
    KHS_JX1 = 32, KHS_JX2, KHS_JX3,    /* kernel_ym_currents (YM_Physics.cu:9-60) */
    KHS_JZ1, KHS_JZ2, KHS_JZ3,
    KHS_UXA, KHS_UZA, KHS_UXB, KHS_UZB,                 /* kernel_update_velocities */
    KHS_SRCPXA, KHS_SRCPZA, KHS_SRCPXB, KHS_SRCPZB,     /* kernel_ym_lorentz */
    KHS_SQ1A, KHS_SQ2A, KHS_SQ3A, KHS_SQ1B, KHS_SQ2B, KHS_SQ3B, /* kernel_ym_precession */
    KHS_ENERGY_DENSITY,                                 /* kernel_ym_energy */
    KHS_N_ARRAY_IDS,
    /* EMHD model (emhdgpu2d): state ids, then derived ids */
    KHS_EM_EX = 64, KHS_EM_EY, KHS_EM_EZ, KHS_EM_BX, KHS_EM_BY, KHS_EM_BZ,
    KHS_EM_AX, KHS_EM_AY, KHS_EM_AZ,
    KHS_EM_N = 76, KHS_EM_PX, KHS_EM_PY, KHS_EM_PZ,          /* fluid 1 */
    KHS_EM_N2 = 82, KHS_EM_PX2, KHS_EM_PY2, KHS_EM_PZ2,      /* fluid 2 (extension) */
    KHS_EM_JX = 96, KHS_EM_JY, KHS_EM_JZ,                    /* kernel_compute_currents */
    KHS_EM_UX, KHS_EM_UZ, KHS_EM_UX2, KHS_EM_UZ2,            /* update_velocities */
    KHS_EM_SX, KHS_EM_SY, KHS_EM_SZ, KHS_EM_SX2, KHS_EM_SY2, KHS_EM_SZ2,  /* kernel_compute_lorentz */
    KHS_EM_ENERGY_DENSITY                                    /* kernel_energy_density */
};

/* Run configuration.  The first block holds the reference's RunConfig keys
 * (ymgpu2d/YM_Config.cuh:6-50) with the same names, defaults and float storage
 * (std::stof semantics, YM_Config.cu:99-123).  The second block holds keys the
 * reference cannot express.  The third block is derived by khs_derive()
 * exactly as main_ym.cu:51-83 and :244-249 derive it. */
typedef struct khs_config {
    /* --- reference keys (example.ini) --- */
    int   k_mode;
    float alpha_YM;
    float perturb_amp;
    int   run_mode;
    float V0;
    float xi_sponge;
    float sigma_sponge;
    float xi_cut;
    float warm_T;
    int   freeze_override;
    int   suppress_kz0;
    float hyp_diff;
    int   kz_suppress_max;
    float eps_override;
    int   kz_suppress_hi;
    float vz_edge_taper;
    int   init_by1_eq;
    int   nz_override;
    int   nx_override;
    float courant_override;
    float lz_override;
    float lx_override;
    float target_tu;
    char  run_tag[128];
    char  seed_profile_file[512];
    /* --- extension keys --- */
    int      n_steps;          /* >0: exact step count, overrides target_tu */
    double   noise_amp;        /* white noise U[-a,a] added to colour-2/3 E and By */
    uint64_t noise_seed;       /* counter-based noise key */
    int      kz_keep_max;      /* >0: keep only |kz| <= K in the 21 filtered arrays */
    int      precision;        /* KHS_FP64 or KHS_FP32 */
    int      energy_stride_override;
    int      export_stride_override;  /* <0 disables snapshots */
    int      fused_sweep;      /* 1 = khs_step runs the fused x+z sweep kernel (YM and EMHD), 0 = the
                                  split x / z sweeps, -1 (default) = the autotuner's choice (fused without
                                  tuning; split when a split-sweep tile shape is forced) */
    int      autotune;         /* 1 (default): time the tile variants on the first step, keep the fastest */
    int      x_segment;        /* 0 = tuned, or force 64 / 128 / 256 cells per x-sweep segment */
    int      z_chunk;          /* 0 = tuned, or force 16 / 32 / 64 rows per z-sweep chunk */
    /* --- model selection and EMHD keys (emhdgpu2d takes k_mode and
     *     perturb_amp from argv, main_coupled.cu:228-229; grid 256x256, dx=1,
     *     dt=0.002 fixed, :236-243) --- */
    int      model;            /* KHS_MODEL_YM (default) or KHS_MODEL_EMHD */
    int      species;          /* EMHD fluids: 1 (reference) or 2 (extension) */
    int      periodic;         /* EMHD: 0 = walls in x (reference), 1 = periodic x (extension) */
    double   charge1, charge2; /* EMHD fluid charges (reference fluid: -1; second fluid: +1) */
    int      emhd_init;        /* 0 = das_kaw shear equilibrium (reference), 1 = random-phase spectrum */
    int      smooth_init;      /* EMHD: one 5-point smoothing of By, Ex after init (reference: 1) */
    int      spectrum_kmax;    /* spectrum init: modes kx, kz = 1..kmax in By */
    double   spectrum_amp;
    uint64_t spectrum_seed;
    double   em_perturb_amp;   /* EMHD das_kaw perturbation (argv[2] as double, default 0.05); set by perturb_amp */
    int      z_groups;         /* z sweep: -1 = autotuned (default), 0 = all variables per lane, 1 = two variable groups on two waves */
    int      xz_variant;       /* fused sweep kernel variant: -1 = autotuned (default), YM 0..14 (13, 14 with warm_T = 0 only), EMHD 0..6 */
    int      x_serial;         /* x sweep phases per variable: -1 = autotuned (default), 0 = all at once, 1 = one at a time */
    int      maxwell_kernel;   /* YM Maxwell: -1 = autotuned (default), 0/2/3 = 64x16/64x4/64x8 LDS tiles, 1 = per cell,
                                  4..10 = z march with 16/32/16/8/8/4/2-row chunks (6, 8: without next-row prefetch);
                                  EMHD: 0 = 64x16 tiles, 1/2/3 = z march with 4/8/16-row chunks */
    int      x_rows;           /* x sweep rows per workgroup: 0 = autotuned (default), 1, 4 or 16 */
    int      z_prefetch;       /* z sweep: -1 = autotuned (default), 0 = load each row when used, 1 / 2 = load one / two rows ahead */
    int      mfma_filter;      /* keep-|kz|<=K filter as MFMA contractions (1, default) or the per-column FP64-sum kernels (0) */
    int      x_lane_cells;     /* x sweep form: -1 = autotuned (default), 0 = LDS-staged segments, 4 or 8 = register form with that many cells per lane (YM) */
    int      snake_order;      /* 1 (default): consecutive step kernels walk the grid in opposite directions (cache reuse); 0: all forward */
    int      zm_fused;         /* YM step: -1 = autotuned (default), 0 = split z sweep and Maxwell, 1/2/3 = one fused kernel with 16/32/64-row chunks */
    int      step_graphs;      /* 1 (default): khs_step replays whole energy strides (even) as one captured HIP graph; 0: eager launches */
    int      step_fusion;      /* YM fused-sweep steps with no column filter (only sponge / xi_cut): Maxwell of a step and
                                  the next step's sweeps as one launch; -1 = the autotuner's choice (default), 0 = off,
                                  1 = on */
    int      mxz_variant;      /* that launch's shape: -1 = autotuned (default), 0..8 (16/32/64-row chunks with
                                  two Maxwell wavefronts per workgroup; 3..5 with four; 6, 7 with four and
                                  chunks sized to fill the device's resident workgroups once / twice; 8 with
                                  two, once) */
    int      packed_shape;     /* khs_batch_run_packed launch shape (read from the batch's first run): -1 = timed on
                                  the first capture (default), 0..7 (see khs_batch_run_packed) */
    /* --- derived (khs_derive) --- */
    int    nx, nz;
    int    periodic_x;         /* run_mode >= 1 */
    int    freeze_az1;
    double lx, lz, dx, dz, courant, dt;
    double alpha, v0, eps;     /* double copies of the float keys */
    double c, eps_0;           /* 1, 1 (main_ym.cu:69) */
    int    total_steps, export_stride, energy_stride;
    double energy_thresh;
} khs_config;

typedef struct khs_ctx khs_ctx;

/* Message of the last failed call on this thread ("" if none). */
const char* khs_last_error(void);
/* Library version string. */
const char* khs_version(void);
/* sizeof(khs_config) as compiled into the library (binding layout check). */
size_t khs_config_sizeof(void);

/* ---------------- configuration (replaces YM_Config.cuh:55 load_config) ---------------- */
int khs_config_defaults(khs_config* cfg);
/* Parse one key/value pair with the reference's parser semantics. Unknown keys
 * return KHS_ERR_CONFIG (the reference only warns; the ini loader below warns). */
int khs_config_set(khs_config* cfg, const char* key, const char* value);
/* Parse an .ini file (YM_Config.cu:15-73), then khs_derive(). */
int khs_load_ini(const char* path, khs_config* cfg);
/* Fill the derived block and run the constraint checks of YM_Config.cu:119-133. */
int khs_derive(khs_config* cfg);
/* Output directory name built from the parameters (main_ym.cu:143-170). */
int khs_output_dir(const khs_config* cfg, char* buf, size_t buflen);

/* ---------------- context ---------------- */
/* Allocate device state for cfg (khs_derive must have run).  Replaces the
 * allocation block of main_ym.cu:86-121. */
int khs_create(const khs_config* cfg, khs_ctx** out);
int khs_destroy(khs_ctx* ctx);
/* Initial state (kernel_ym_init, YM_Init.cu:32-257, evaluated on the host in
 * the same float arithmetic) plus the optional eigenmode seed file
 * (main_ym.cu:178-217).  seed_path NULL uses cfg->seed_profile_file. */
int khs_init(khs_ctx* ctx, const char* seed_path);

/* ---------------- state I/O ---------------- */
int khs_array_id(const char* name);
int khs_get_array(khs_ctx* ctx, int id, double* host, size_t n);
int khs_set_array(khs_ctx* ctx, int id, const double* host, size_t n);
/* Read a state array from one of the step buffers: the committed state, the
 * x-swept stage (fluid ids 12..23 only) or the next state being built by the
 * staged operators below.  Lets a caller inspect the result of each operator,
 * as the reference's device arrays can be read between kernel launches. */
#define KHS_BUF_CURRENT 0
#define KHS_BUF_STAGE   1
#define KHS_BUF_NEXT    2
int khs_get_array_buf(khs_ctx* ctx, int buf, int id, double* host, size_t n);

/* ---------------- operators (reference call boundaries) ----------------
 * A step is staged: x-sweeps read the committed state and write a stage
 * buffer; z-sweeps read the stage and write the next state; khs_maxwell reads
 * the next fluid and committed fields and writes the next fields; khs_filters
 * acts on the next state; khs_commit makes it current.  Calling them in the
 * order of main_ym.cu:322-455 reproduces khs_step.  Velocities and Lorentz /
 * precession sources are evaluated on the fly from the committed state, which
 * is the value the reference holds in its d_src / d_sQ arrays at that point
 * (they are computed from the end-of-step state, main_ym.cu:450-455). */
int khs_fct_x_sweep(khs_ctx* ctx, int species);   /* ym_fct_x_sweep[_periodic], YM_FCT.cu:30-91 */
int khs_fct_z_sweep(khs_ctx* ctx, int species);   /* ym_fct_z_sweep, YM_FCT.cu:93-135; + enforce_boundaries on the fluid */
int khs_maxwell(khs_ctx* ctx);                   /* wall Ez^a = 0 + currents + Ampere + potential + Faraday */
/* kernel_update_velocities (YM_Fluid.cu:3-11; EMHD update_velocities): ux =
 * px / max(n, 0.05) and uz = pz / max(n, 0.05) of species s of the committed
 * state, into host buffers of nx*nz doubles.  The sweeps evaluate the same
 * values on the fly; this exports them. */
int khs_fluid_velocities(khs_ctx* ctx, int species, double* ux, double* uz, size_t n);
/* kernel_ym_lorentz / kernel_ym_precession (YM_Fluid.cu:32-108): the sources
 * that the next step's sweeps add for species s (px and pz momentum sources,
 * then the three colour precession rates in sQ[0..3n)), evaluated on the
 * committed state.  EMHD: kernel_compute_lorentz's (sx, sy, sz) into src_px,
 * src_pz and sQ[0..n) (sy).  Host buffers of n = nx*nz doubles (sQ 3n). */
int khs_sources(khs_ctx* ctx, int species, double* src_px, double* src_pz, double* sQ, size_t n);
/* kernel_enforce_boundaries (YM_Fluid.cu:13-30; EMHD main_coupled.cu:48-63)
 * on the committed state, in place: density floor 0.05; with walls in x,
 * px = 0 and Ez^a = 0 (EMHD: Ey, Ez, Bx = 0) in the wall columns.  khs_step
 * and khs_fct_z_sweep / khs_maxwell apply the same conditions as part of
 * their outputs; this operator serves callers that write arrays directly. */
int khs_enforce_boundaries(khs_ctx* ctx);
int khs_filters(khs_ctx* ctx);                    /* sponge, xi_cut, z-mean, kz bands, hyperdiffusion, colour-1 zeroing */
int khs_commit(khs_ctx* ctx);

/* Fused step: nsteps full steps.  Energy is checked every energy_stride steps
 * (counted from the context's step counter) with the reference's halt rule
 * (main_ym.cu:457-475); *halted (may be NULL) is set to 1 and stepping stops
 * on NaN/Inf or E/E0 > energy_thresh. */
int khs_step(khs_ctx* ctx, int nsteps, int* halted);
/* Current step counter and simulation time. */
int khs_get_step(khs_ctx* ctx, int* step, double* time);

/* ---------------- diagnostics ---------------- */
/* Total energy (YM_Energy.cu:3-32) as a compensated FP64 device reduction. */
int khs_energy(khs_ctx* ctx, double* E);
/* The energy that the last energy check of khs_step (main_ym.cu:457-479) or
 * of a batch segment used, and the step it was taken at (-1: no check yet).
 * Where the steps before a check run as Maxwell + next-sweep launches, the
 * check's energy is summed by the last Maxwell launch itself, so this is the
 * value the halting rule saw.  No counterpart in the reference. */
int khs_last_energy(khs_ctx* ctx, double* E, int* step);
/* Allocation-free energy reduction on caller-owned device arrays.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc), which compute_ym_energy (ymgpu2d/YM_Energy.cuh:10-12)
 * cannot serve because its caller must supply an NX*NZ per-cell buffer.
 * arrays[id] are device pointers in the order of the enum khs_array ids, layout
 * x + z*nx, double for KHS_FP64 and float for KHS_FP32; the 15 arrays the
 * energy reads (Ex, Ez, By of every colour, n, px, pz of both species) must
 * be non-null.  The kernels are enqueued on `stream` (a hipStream_t, NULL =
 * the null stream) and write the compensated FP64 sum of YM_Energy.cu:3-32
 * to out_dev[0] (device memory); the call neither allocates nor waits, so it
 * may be captured into a HIP graph.  The reduction's partial sums live in a
 * fixed 16 KiB device array of the library (1024 blocks, any grid up to the
 * library's 2^29 - 2 cells), one per GPU, shared by every call on that GPU:
 * calls that overlap on one GPU (other streams, other threads) race. */
int khs_energy_noalloc(const void* const arrays[KHS_N_STATE], int nx, int nz, int precision, void* stream,
                       double* out_dev);
/* Initial energy E0 recorded by khs_init. */
int khs_energy0(khs_ctx* ctx, double* E0);
/* Gauss-law residual (analysis/gauss_check.py:40-115 formula, computed on the
 * device in FP64).  out[0..2] = rms G^a with rho = +nQ (reference sign),
 * out[3..5] = rms G^a with rho = -nQ (continuity-consistent sign),
 * out[6] = Etot_scale = rms(sum|E|)/dx. */
int khs_gauss(khs_ctx* ctx, double out[7]);
/* The same residual as gauss_check.py reports it (:95-107): rms[a] = rms G^a
 * with rho = +nQ, rel[a] = rms[a] / (rms(sum|E|)/dx). */
int khs_gauss_residual(khs_ctx* ctx, double rms[3], double rel[3]);
/* Circular amplitude of the colour-2/3 potential at z mode kz
 * (analysis/remote_timeseries.py:62-77): F(x) = sum_z (Az2 + i Az3)
 * e^{-2 pi i kz z/NZ}; out[0] = mean_x |F| / NZ, out[1] = the same for
 * Az2 - i Az3.  FP64 device reduction.  YM model only. */
int khs_mode_amplitude(khs_ctx* ctx, int kz, double out[2]);

/* 2-D Fourier coefficients of a state array (id of the run's model) at
 * kx, kz = 1..kmax (kmax <= 32): out[2 (kz-1) kmax + 2 (kx-1) + {0, 1}] =
 * {Re, Im} of sum_{x,z} f e^{-2 pi i (kx x / NX + kz z / NZ)}.  The FP32 form
 * contracts z and then x on MFMA (v_mfma_f32_16x16x4_f32); the FP64 form
 * with plain FP64 sums.  Used for the space-time dispersion of wave runs. */
int khs_mode_spectrum(khs_ctx* ctx, int id, int kmax, double* out);

/* ---------------- growth-rate fits (host only, no context) ---------------- */
/* Max-R^2 window fit of log(amp) vs t (analysis/batch_analyze.py:44-69).
 * t_min / t_max: NaN = unbounded.  out = {gamma, t0, t1, r2}; fewer than 4
 * samples give {0, NaN, NaN, 0}, no window with slope >= 1e-6 gives r2 = -1e9. */
int khs_fit_growth_max_r2(const double* t, const double* amp, int n, double t_min, double t_max,
                          double min_win_frac, double out[4]);
/* Plateau fit (batch_analyze.py:90-152).  out = {gamma, t0, t1, duration,
 * has_plateau (1/0)}; without a plateau gamma, t0, t1, duration are NaN. */
int khs_fit_growth_plateau(const double* t, const double* amp, int n, double win_tu,
                           double min_plateau_tu, double flat_tol, double out[5]);

/* ---------------- output files (byte-compatible with main_ym.cu) ---------------- */
int khs_write_snapshot(khs_ctx* ctx, const char* dir, int step);      /* ym_%06d.csv, :287-306 */
int khs_energy_csv_open(khs_ctx* ctx, const char* path);              /* header + "0,0,E0,1.0", :252-255 */
int khs_energy_csv_row(khs_ctx* ctx, int step, double E);            /* step,time,E,E/E0, :462 */
int khs_energy_csv_close(khs_ctx* ctx);

/* ---------------- binary checkpoint / restart ---------------- */
/* No counterpart in the reference (its runs restart only from a seed file).
 * khs_checkpoint_write stores the committed state: an 80-byte header
 * {char magic[8] = "KHSCKPT", uint32 version = 1, int32 model, nx, nz, fp32,
 * species, periodic_x, n_arrays, int64 step, double E0, dt, dx, dz}, then
 * n_arrays x nx*nz FP64 arrays (state slots 0..23 in the order of enum
 * khs_array; the FP32 form widens exactly), then a uint64 FNV-1a hash of the
 * array bytes.  The file is written to path + ".tmp" and renamed.
 * khs_checkpoint_read loads such a file into a context created from the same
 * configuration (model, grid, precision, species, x boundary and dt / dx / dz
 * must match: KHS_ERR_CONFIG otherwise; KHS_ERR_IO for a short, corrupt or
 * foreign file) and restores the step count and E0, so stepping on gives the
 * bits of an uninterrupted run.  Both need a committed step (KHS_ERR_STATE). */
int khs_checkpoint_write(khs_ctx* ctx, const char* path);
int khs_checkpoint_read(khs_ctx* ctx, const char* path);

/* ---------------- timing ---------------- */
int khs_sync(khs_ctx* ctx);
/* Enable per-kernel HIP-event timing on the context's stream.  While enabled,
 * every launch of the step kernels is bracketed by events; khs_kernel_times
 * returns, per kernel slot, the summed milliseconds and the launch count since
 * the last reset.  Slot names come from khs_kernel_name(). */
int khs_timing_enable(khs_ctx* ctx, int on);
int khs_timing_reset(khs_ctx* ctx);
int khs_kernel_times(khs_ctx* ctx, double* ms, int* launches, int nslots);
const char* khs_kernel_name(int slot);
/* Tile shapes in use: out = {x-sweep segment width (cells), x-sweep rows per
 * workgroup, z-sweep chunk height (rows), z-sweep next-row prefetch (0/1),
 * YM Maxwell kernel (0/2/3 = 64x16/64x4/64x8 LDS tiles, 1 = per cell, 4..10 = z march), x-sweep variables one at a
 * time (0/1), fused-sweep variant (0: one wave per 64 columns; 1/2/3: four
 * waves per 128 columns with 16/32/64-row chunks; 4/5/6: register x sweep in
 * the z march, 32/64/128-row chunks; 7/8/9: the same with the variables on two
 *  waves, 16/32/64-row chunks, two column tiles per workgroup; 10 / 11 / 12: one tile per workgroup, 32 / 8 / 4-row chunks; EMHD: 0..3 = 16/32/64/8-row chunks, 4..6 = two waves per species, 8/16/32-row chunks), z-sweep variables split
 * over two waves (0/1), x-sweep cells per lane of the register form (0 = LDS
 * form; the register form reports its 62*C-cell wave segment, 1 row, and 1
 * for variables one at a time), fused z sweep + Maxwell kernel (0 = split
 * kernels, 1/2/3 = fused with 16/32/64-row chunks), fused x+z sweep in use (0/1), Maxwell and the next
 * step's sweeps in one launch (0/1), that launch's variant (see mxz_variant)}; chosen by the autotuner on the
 * first step unless autotune = 0. */
int khs_get_tuning(khs_ctx* ctx, int out[13]);

/* ---------------- batches of runs (parameter sweeps) ---------------- */
/* No counterpart in the reference, which runs one configuration per process
 * (scripts/run_multigpu_abi_template.sh).  A batch owns one context per config;
 * khs_batch_run advances every active run by nsteps: `segment` (even) steps of
 * all runs plus their energy reductions form one captured HIP graph, the runs
 * spread over `lanes` streams so that small grids execute concurrently.  The
 * halting rule (NaN/Inf, E/E0 > threshold) is applied after every segment
 * rather than every energy_stride steps.  steps_done[i] / halted[i] (n
 * entries each, may be NULL) report each run's step count and halt flag. */
typedef struct khs_batch khs_batch;
int khs_batch_create(const khs_config* cfgs, int n, int lanes, int segment, khs_batch** out);
int khs_batch_destroy(khs_batch* b);
int khs_batch_init(khs_batch* b);
int khs_batch_run(khs_batch* b, int nsteps, int* steps_done, int* halted);
/* The same segments with the runs packed into shared launches: every step of
 * all active runs is one fused x+z sweep launch, one Maxwell launch and one
 * filter launch whose tiles are the runs' tiles one run after another (each
 * workgroup steps its run with that run's own state and parameters, so the
 * results are the bits of stepping each run alone), the segment and the
 * runs' energy reductions captured as one HIP graph.  Small grids then share
 * the device instead of waiting on launch latency.  All runs must use the
 * YM model and one precision, and filter only with xi_cut / sponge /
 * suppress_kz0 (KHS_ERR_CONFIG otherwise).  Launch shapes (chunk rows of the
 * sweep, of the Maxwell march; two sweep tiles per workgroup unless noted):
 * 0: 16, 8; 1: 32, 4; 2: 32, 8; 3: 16, 4; 4: 32, 2; 5: 16, 2; 6: 32, 4 and
 * 7: 8, 4 with one sweep tile per workgroup; timed on the first capture
 * unless the first run's packed_shape forces one. */
int khs_batch_run_packed(khs_batch* b, int nsteps, int* steps_done, int* halted);
/* The same stepping without graphs: each run in turn through khs_step
 * (halting every energy_stride steps), for comparison. */
int khs_batch_run_eager(khs_batch* b, int nsteps, int* steps_done, int* halted);
/* The runs stepped concurrently: dealt to `threads` host threads by longest
 * processing time first, each thread driving its runs with khs_step (whole
 * energy strides replayed as captured graphs, the reference's halting rule
 * at every energy check) on the run's own stream, so that small,
 * latency-bound runs overlap with each other and with large ones.  Step
 * limits (khs_batch_set_step_limit) apply. */
int khs_batch_run_concurrent(khs_batch* b, int nsteps, int threads, int* steps_done, int* halted);
int khs_batch_get_array(khs_batch* b, int run, int id, double* host, size_t n);
int khs_batch_set_array(khs_batch* b, int run, int id, const double* host, size_t n);
int khs_batch_energy(khs_batch* b, int run, double* E, double* E0);
/* Step limit of one run (max_step <= 0: none, the default): khs_batch_run
 * stops the run there without marking it halted. */
int khs_batch_set_step_limit(khs_batch* b, int run, int max_step);
/* khs_mode_amplitude of one run of the batch (campaign growth-rate sampling) */
int khs_batch_mode_amplitude(khs_batch* b, int run, int kz, double out[2]);

#ifdef __cplusplus
}
#endif

#endif /* KHS_H */
