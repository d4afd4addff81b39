// khs_context.cpp — device context, step orchestration and the C ABI of libkhs.
//
// Memory layout in HBM: one allocation holding two full state sets (24 arrays
// each; a frozen Az1 is shared by both) and a 12-array stage buffer for the
// x-swept fluid.  Each array is nx*nz elements of the run's precision in the
// reference layout (x contiguous), padded to 256 bytes.  A step ping-pongs
// between the state sets so that no kernel updates a stencil input in place:
//
//   k_fct_x   : state[cur] -> stage                (both species, one launch)
//   k_fct_z   : stage + state[cur] -> state[nxt].fluid
//   k_maxwell : state[nxt].fluid + state[cur].fields -> state[nxt].fields
//   filters   : state[nxt] in place (column-local / cell-local only)
//   commit    : cur <-> nxt
//
// This replaces main_ym.cu:322-455 (about 116 launches, 44 device syncs and 32
// device copies per step) with three launches and no host synchronisation;
// the host only synchronises for the energy check every energy_stride steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <mutex>
#include <map>
#include <chrono>
#include <functional>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <memory>
#include <sstream>
#include <thread>
#include <string>
#include <type_traits>
#include <vector>

#include "khs_host.h"
#include "khs_launch.h"

namespace khs {

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            throw Error(KHS_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));   \
    } while (0)

enum Slot { K_FCT_X = 0, K_FCT_Z, K_FCT_XZ, K_MAXWELL, K_FILTERS, K_ENERGY, K_FCT_ZM, K_MXZ, K_NSLOTS };
static const char* kSlotNames[K_NSLOTS] = {"k_fct_x", "k_fct_z", "k_fct_xz", "k_maxwell", "filters", "energy",
                                           "k_fct_zm", "k_mxz"};

static const int kFilt15[15] = {S_BY1, S_EX1, S_EZ1, S_BY2, S_BY3, S_EX2, S_EX3, S_EZ2, S_EZ3,
                                S_AZ2, S_AZ3, S_Q2A, S_Q3A, S_Q2B, S_Q3B};
static const int kFluid6[6] = {S_PZA, S_PZB, S_PXA, S_PXB, S_NA, S_NB};
static const int kCol23[12] = {S_BY2, S_BY3, S_EX2, S_EX3, S_EZ2, S_EZ3, S_AZ2, S_AZ3,
                               S_Q2A, S_Q3A, S_Q2B, S_Q3B};
static const int kQ23[4] = {S_Q2A, S_Q3A, S_Q2B, S_Q3B};

struct ModeTable {
    void* c = nullptr;   // [nmodes][nz] of float or double
    void* s = nullptr;
    void* w = nullptr;   // [nmodes]
    int nmodes = 0;
    bool is_float = false;
};

// process-wide cache of autotuned kernel choices, keyed by Ctx::tune_key()
static std::map<std::vector<long>, std::array<int, 8>>& tune_cache() {
    static std::map<std::vector<long>, std::array<int, 8>> m;
    return m;
}
static std::mutex& tune_cache_mutex() {
    static std::mutex m;
    return m;
}
// Stream capture and graph instantiation are serialised over the process:
// host threads that drive separate contexts (the concurrent batch runner,
// callers with one context per thread) then never have two captures or
// instantiations in flight inside the HIP runtime at once.  Launching and
// synchronising captured graphs stays concurrent.
static std::mutex& capture_mutex() {
    static std::mutex m;
    return m;
}

struct Ctx {
    khs_config cfg{};
    KParams kp{};
    int nx = 0, nz = 0;
    size_t N = 0, stride = 0, esz = 8;
    bool fp32 = false;
    bool emhd = false;      // emhdgpu2d model (fields 0..8, fluids at 12 / 18)
    int nsp = 2;            // fluid species of the model

    char* dmem = nullptr;
    void* st[2][S_COUNT] = {};
    void* stage[2 * F_COUNT] = {};
    int cur = 0;

    // filters
    double* d_decay = nullptr;
    int* d_action = nullptr;
    bool cell_filters = false;
    ModeTable t_zmean, t_lo, t_hi, t_keep;
    double* d_coef = nullptr;
    void* d_keep_tab = nullptr;      // MFMA keep filter table (nz x 16, run precision)
    // mode spectrum (allocated on first use)
    char* spec_mem = nullptr;
    void *spec_zt = nullptr, *spec_xt = nullptr, *spec_C = nullptr, *spec_D = nullptr, *spec_P = nullptr;
    int spec_K = 0;

    // reductions / derived
    double* d_partials = nullptr;
    double* d_red = nullptr;
    double* h_red = nullptr;   // pinned
    double* d_derived = nullptr;

    hipStream_t stream = nullptr;

    // timing
    bool timing = false;
    struct Pending { int slot; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> evpool;
    double ms[K_NSLOTS] = {};
    int cnt[K_NSLOTS] = {};

    // tile variants (autotuned on the first step)
    int xvar = 2, zvar = 1, mvar = 0, xzvar = 7;
    bool fused = false;                 // khs_step runs the fused x+z sweep
    int zmvar = 0;                            // 0: split z sweep + Maxwell; 1..3: fused kernel
    // Maxwell of one step and the sweeps of the next in one launch (k_mxz):
    // khs_step runs a segment of L steps as sweeps, L - 1 of those launches,
    // then Maxwell and the filters
    bool mxz = false;
    int mxzvar = 1;
    // per-wavefront energy sums of the last Maxwell of an energy stride
    double* d_epart = nullptr;
    int n_epart = 0;
    bool energy_ready = false;   // d_red holds the energy of the state at step energy_step
    long energy_step = -1;       // (summed by that step's Maxwell)
    double last_check_E = 0.0;   // the energy of khs_step's last check, and its step
    long last_check_step = -1;
    // one energy stride of steps + the energy reduction, captured once
    hipGraphExec_t sgraph[2] = {nullptr, nullptr};   // by the ping-pong buffer a stride starts from
    bool tuned = false;

    // step state
    long step = 0;
    double E0 = 0.0;
    int staged_x = 0, staged_z = 0;
    bool maxwell_done = false, filters_done = false;
    std::ofstream ecsv;

    ~Ctx();
    template <typename T> State<T> state(int k) const {
        State<T> s;
        for (int i = 0; i < S_COUNT; ++i) s.a[i] = static_cast<T*>(st[k][i]);
        return s;
    }
    // the 12 fields of buffer set fk with the 12 fluid arrays of set lk
    template <typename T> State<T> mixed(int fk, int lk) const {
        State<T> s;
        for (int i = 0; i < S_COUNT; ++i) s.a[i] = static_cast<T*>(st[i < S_NA ? fk : lk][i]);
        return s;
    }
    // k_mxz needs a step whose only filters are per column (sponge, xi_cut)
    // and no warm-closure terms
    bool mxz_eligible() const {
        const khs_config& c = cfg;
        const bool column = c.suppress_kz0 || c.kz_suppress_max >= 1 || c.kz_suppress_hi > c.kz_suppress_max + 1 ||
                            c.kz_keep_max > 0 || c.hyp_diff > 0.0f;
        // k_mxz drops the factors c^2 and 1/eps0 (exact only for c = eps0 = 1,
        // the reference's fixed values)
        return !emhd && fused && !column && kp.warm_T <= 0.0 && kp.c2 == 1.0 && kp.eps0 == 1.0;
    }
    template <typename T> Stage<T> stg() const {
        Stage<T> s;
        for (int i = 0; i < 2 * F_COUNT; ++i) s.a[i] = static_cast<T*>(stage[i]);
        return s;
    }
    template <typename T> ArrayList<T> list(int k, const int* ids, int n, bool skip_c1) const {
        ArrayList<T> l{};
        l.n = 0;
        for (int i = 0; i < n; ++i) {
            int id = ids[i];
            if (skip_c1 && (id == S_BY1 || id == S_EX1 || id == S_EZ1)) continue;
            l.a[l.n++] = static_cast<T*>(st[k][id]);
        }
        return l;
    }

    hipEvent_t get_event() {
        if (!evpool.empty()) { hipEvent_t e = evpool.back(); evpool.pop_back(); return e; }
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        return e;
    }
    template <typename F> void timed(int slot, F&& f) {
        if (!timing) { f(); return; }
        hipEvent_t a = get_event(), b = get_event();
        HIPCHK(hipEventRecord(a, stream));
        f();
        HIPCHK(hipEventRecord(b, stream));
        pending.push_back({slot, a, b});
    }
    void harvest() {
        if (pending.empty()) return;
        HIPCHK(hipStreamSynchronize(stream));
        for (auto& p : pending) {
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, p.a, p.b));
            ms[p.slot] += t;
            cnt[p.slot] += 1;
            evpool.push_back(p.a);
            evpool.push_back(p.b);
        }
        pending.clear();
    }
    void check_launch() {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) throw Error(KHS_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    }

    // ---- stages ----
    template <typename T> void fct_x(int sp0, int ns, int v) {
        if (emhd) launch_fct_x<T, EMModel>(state<T>(cur), stg<T>(), kp, sp0, ns, v, stream);
        else launch_fct_x<T, YMModel>(state<T>(cur), stg<T>(), kp, sp0, ns, v, stream);
    }
    template <typename T> void fct_z(int sp0, int ns, int v) {
        if (emhd) launch_fct_z<T, EMModel>(state<T>(cur), stg<T>(), state<T>(1 - cur), kp, sp0, ns, v, stream);
        else launch_fct_z<T, YMModel>(state<T>(cur), stg<T>(), state<T>(1 - cur), kp, sp0, ns, v, stream);
    }
    // Direction of the step kernels' tile walk: x and Maxwell one way, the z
    // sweep between them the other, flipping every step, so each kernel
    // starts where its predecessor ended.
    void set_order(int stage) {
        kp.rev = cfg.snake_order ? (int)((step + (stage == 1)) & 1) : 0;
    }
    template <typename T> void do_x(int sp0, int ns) {
        set_order(0);
        timed(K_FCT_X, [&] { fct_x<T>(sp0, ns, xvar); });
        check_launch();
    }
    template <typename T> void do_z(int sp0, int ns) {
        set_order(1);
        timed(K_FCT_Z, [&] { fct_z<T>(sp0, ns, zvar); });
        check_launch();
    }
    // Time every tile variant of the two sweeps on this grid (3 launches each
    // after a warm-up, two passes over the list) and keep the fastest.  Runs
    // before the first sweep of the context, when the stage and the
    // next-state fluid are scratch; all variants compute identical bits.
    template <typename T> int pick(const std::vector<int>& cand, const std::function<void(int)>& launch,
                                   float* best_ms = nullptr) {
        hipEvent_t a = get_event(), b = get_event();
        float best = 1e30f;
        int bi = cand[0];
        // two passes over the candidates, each variant's best pass counts: a
        // single pass favours or penalises variants by the cache state their
        // predecessor in the list leaves
        std::vector<float> tv(cand.size(), 1e30f);
        for (int pass = 0; pass < (cand.size() > 1 ? 2 : 1); ++pass)
            for (size_t j = 0; j < cand.size(); ++j) {
                const int v = cand[j];
                launch(v);
                HIPCHK(hipEventRecord(a, stream));
                for (int r = 0; r < 3; ++r) launch(v);
                HIPCHK(hipEventRecord(b, stream));
                HIPCHK(hipEventSynchronize(b));
                float t = 0.f;
                HIPCHK(hipEventElapsedTime(&t, a, b));
                tv[j] = std::min(tv[j], t);
            }
        for (size_t j = 0; j < cand.size(); ++j)
            if (tv[j] < best) { best = tv[j]; bi = cand[j]; }
        check_launch();
        evpool.push_back(a);
        evpool.push_back(b);
        if (best_ms) *best_ms = best;
        return bi;
    }
    // the kernel-shape keys and everything else the timing depends on
    std::vector<long> tune_key() const {
        return {nx, nz, fp32, emhd, nsp, kp.periodic_x, cell_filters, cfg.fused_sweep, cfg.x_segment, cfg.x_rows,
                cfg.x_serial, cfg.x_lane_cells, cfg.z_chunk, cfg.z_prefetch, cfg.z_groups, cfg.maxwell_kernel,
                cfg.zm_fused, cfg.xz_variant, cfg.step_fusion, cfg.mxz_variant};
    }
    template <typename T> void do_autotune() {
        tuned = true;
        const int fx = cfg.x_segment == 64 ? 0 : cfg.x_segment == 128 ? 1 : cfg.x_segment == 256 ? 2 : -1;
        const int fz = cfg.z_chunk == 16 ? 0 : cfg.z_chunk == 32 ? 1 : cfg.z_chunk == 64 ? 2 : -1;
        const int fp = cfg.z_prefetch;               // -1 = tuned
        std::vector<int> xc, zc;
        const int fr = cfg.x_rows == 1 ? 0 : cfg.x_rows == 4 ? 1 : cfg.x_rows == 16 ? 2 : -1;
        const int fv = cfg.x_serial;             // -1 = tuned
        const int fl = cfg.x_lane_cells;         // -1 = tuned, 0 = LDS form, 4 / 8 = register form
        const bool lds_keys = fx >= 0 || fr >= 0 || fv >= 0;
        for (int v = 0; v < 18; ++v)
            if ((fl <= 0) && (fx < 0 || v % 3 == fx) && (fr < 0 || (v / 3) % 3 == fr) && (fv < 0 || v / 9 == fv))
                xc.push_back(v);
        if (!emhd && !lds_keys) {
            if (fl < 0 || fl == 4) xc.push_back(18);
            if (fl < 0 || fl == 8) xc.push_back(19);
        }
        const int fg = cfg.z_groups;             // -1 = tuned
        for (int v = 0; v < kNumZVariants; ++v)
            if ((fz < 0 || v % 3 == fz) && (fp < 0 || z_prefetch_of(v) == fp) && (fg < 0 || z_groups_of(v) == fg))
                zc.push_back(v);
        // without tuning: the forced shape, else 256-cell segments and 32-row
        // chunks with prefetch in FP32 (registers to spare), without in FP64
        xvar = ((fx >= 0) ? fx : 2) + 3 * ((fr >= 0) ? fr : 1) + 9 * ((fv >= 0) ? fv : 1);
        if (!emhd && fl > 0) xvar = fl == 4 ? 18 : 19;
        zvar = ((fz >= 0) ? fz : 1) + 3 * ((fp >= 0) ? fp : (fp32 ? 1 : 0)) + 6 * ((fg >= 0) ? fg : 0);
        if (fp == 2) zvar = 12 + ((fz >= 0) ? fz : 1);
        mvar = (cfg.maxwell_kernel >= 0) ? cfg.maxwell_kernel : 0;
        // fused x+z sweep (YM): forced by fused_sweep = 1; with -1 it is the
        // untuned default and the autotuner's candidate against the split
        // sweeps, unless a split-sweep shape is forced
        const bool split_keys = lds_keys || fl >= 0 || fz >= 0 || fp >= 0 || fg >= 0 || cfg.zm_fused > 0;
        const bool try_fused = cfg.fused_sweep < 0 && !split_keys;
        fused = cfg.fused_sweep > 0 || try_fused;
        xzvar = (cfg.xz_variant >= 0) ? cfg.xz_variant : (emhd ? 5 : 7);
        mxzvar = (cfg.mxz_variant >= 0) ? cfg.mxz_variant : 1;
        mxz = cfg.step_fusion != 0 && mxz_eligible();
        if (!cfg.autotune) return;
        // contexts of the same shape in one process (sweeps, campaigns) reuse
        // the first one's choice: every variant computes the same bits
        const std::vector<long> key = tune_key();
        {
            std::lock_guard<std::mutex> lk(tune_cache_mutex());
            auto it = tune_cache().find(key);
            if (it != tune_cache().end()) {
                xvar = it->second[0]; zvar = it->second[1]; mvar = it->second[2];
                xzvar = it->second[3]; zmvar = it->second[4]; fused = it->second[5] != 0;
                mxzvar = it->second[6]; mxz = it->second[7] != 0;
                return;
            }
        }
        struct Store {
            Ctx* c;
            std::vector<long> key;
            ~Store() {
                std::lock_guard<std::mutex> lk(tune_cache_mutex());
                tune_cache()[key] = {c->xvar, c->zvar, c->mvar, c->xzvar, c->zmvar, c->fused ? 1 : 0, c->mxzvar,
                                     c->mxz ? 1 : 0};
            }
        } store{this, key};
        float tx = 0.f, tz = 0.f, txz = 0.f;
        if (xc.size() > 1 || try_fused) xvar = pick<T>(xc, [&](int v) { fct_x<T>(0, nsp, v); }, &tx);
        fct_x<T>(0, nsp, xvar);
        if (zc.size() > 1 || try_fused) zvar = pick<T>(zc, [&](int v) { fct_z<T>(0, nsp, v); }, &tz);
        if (fused) {
            std::vector<int> xzc;
            for (int v = 0; v < (emhd ? kNumXZEmVariants : kNumXZVariants); ++v)
                if ((cfg.xz_variant < 0 && (emhd || v < kFirstColdXZVariant || kp.warm_T <= 0.0)) ||
                    v == cfg.xz_variant)
                    xzc.push_back(v);
            if (xzc.size() > 1 || try_fused)
                xzvar = pick<T>(xzc, [&](int v) { fct_xz<T>(v); }, &txz);
            if (try_fused) fused = txz < tx + tz;
        }
        mxz = cfg.step_fusion != 0 && mxz_eligible();
        zmvar = (!emhd && cfg.zm_fused > 0) ? cfg.zm_fused : 0;
        if (!emhd && cfg.zm_fused < 0 && !fused) {
            // split z sweep + (tuned) Maxwell against the fused kernel's chunk heights
            if (cfg.maxwell_kernel < 0) {
                std::vector<int> mc;
                for (int v = 0; v < kNumMaxwellVariants; ++v) mc.push_back(v);
                mvar = pick<T>(mc, [&](int v) {
                    launch_maxwell<T>(state<T>(cur), state<T>(1 - cur), kp, cell_filters ? d_decay : nullptr,
                                      cell_filters ? d_action : nullptr, v, stream);
                });
            }
            std::vector<int> zmc;
            for (int v = 0; v < kNumZMVariants; ++v) zmc.push_back(v);
            zmvar = pick<T>(zmc, [&](int v) { zm_launch<T>(v); });
            return;
        }
        if (emhd && cfg.maxwell_kernel < 0) {
            std::vector<int> mc;
            for (int v = 0; v < kNumEmMaxwellVariants; ++v) mc.push_back(v);
            // writes the next state's fields from the current state: scratch here
            mvar = pick<T>(mc, [&](int v) { launch_emhd_maxwell<T>(state<T>(cur), state<T>(1 - cur), kp, v, stream); });
        }
        if (!emhd) {
            std::vector<int> mc;
            if (cfg.maxwell_kernel < 0)
                for (int v = 0; v < kNumMaxwellVariants; ++v) mc.push_back(v);
            else
                mc.push_back(mvar);
            // writes the next state's fields from the current state: scratch here
            float tmx = 0.f;
            if (mc.size() > 1 || mxz)
                mvar = pick<T>(mc, [&](int v) {
                    launch_maxwell<T>(state<T>(cur), state<T>(1 - cur), kp, cell_filters ? d_decay : nullptr,
                                      cell_filters ? d_action : nullptr, v, stream);
                }, &tmx);
            if (mxz) {
                // Maxwell + next sweeps in one launch against the two kernels,
                // timed on the current state into the next-state set (scratch)
                std::vector<int> vc;
                for (int v = 0; v < kNumMxzVariants; ++v)
                    if (cfg.mxz_variant < 0 || v == cfg.mxz_variant) vc.push_back(v);
                float tmxz = 0.f;
                mxzvar = pick<T>(vc, [&](int v) {
                    launch_maxwell_sweep<T>(state<T>(cur), state<T>(1 - cur), kp, cell_filters ? d_decay : nullptr,
                                            cell_filters ? d_action : nullptr, v, stream);
                }, &tmxz);
                if (cfg.step_fusion < 0) {
                    if (txz <= 0.f) pick<T>({xzvar}, [&](int v) { fct_xz<T>(v); }, &txz);
                    mxz = tmxz < txz + tmx;
                }
            }
        }
    }
    void autotune() {
        if (tuned) return;
        fp32 ? do_autotune<float>() : do_autotune<double>();
    }
    // fused x+z sweep of every species: current state -> next state
    template <typename T> void fct_xz(int v) {
        if (emhd) launch_fct_xz_em<T>(state<T>(cur), state<T>(1 - cur), kp, nsp, v, stream);
        else launch_fct_xz<T>(state<T>(cur), state<T>(1 - cur), kp, v, stream);
    }
    template <typename T> void do_xz() {
        timed(K_FCT_XZ, [&] { fct_xz<T>(xzvar); });
        check_launch();
    }
    // Maxwell of the step whose fields are in set X and swept fluid in set
    // Y, then the next step's sweeps: fields -> Y, fluid -> X
    template <typename T> void mxz_launch(int X, int Y, int v) {
        launch_maxwell_sweep<T>(mixed<T>(X, Y), mixed<T>(Y, X), kp, cell_filters ? d_decay : nullptr,
                                cell_filters ? d_action : nullptr, v, stream);
    }
    // L >= 2 steps from set cur: the sweeps of the first step, L - 1 launches
    // of Maxwell + the next step's sweeps, then Maxwell, the filters and the
    // commit of the last step.  Between the launches the state is one step
    // apart (fields of step n in X, fluid after the sweeps of step n in Y);
    // the charges' filter of each inner step is applied by the next launch as
    // it reads them, so nothing but the last step's filters is left to do.
    template <typename T> void do_mxz_steps(int L, bool energy_next) {
        set_order(0);
        do_xz<T>();
        int X = cur, Y = 1 - cur;
        for (int j = 1; j < L; ++j) {
            step += 1;
            set_order(0);
            timed(K_MXZ, [&] { mxz_launch<T>(X, Y, mxzvar); });
            check_launch();
            std::swap(X, Y);
        }
        cur = X;                  // fields of the last step's start; its swept fluid in 1 - cur
        // an energy check follows: the last Maxwell sums the new state's
        // energy as it writes it (the charge filter after it leaves the
        // energy unchanged), instead of a separate pass over 15 arrays
        int nw = 0;
        // (the reduction's first stage leaves one pair per 4 kRedThreads
        // wavefront sums in d_partials, which holds kRedBlocks * 8 pairs)
        if (energy_next && d_epart && (n_epart + 4 * kRedThreads - 1) / (4 * kRedThreads) <= kRedBlocks * 8) {
            set_order(2);
            timed(K_MAXWELL, [&] {
                nw = launch_maxwell_energy<T>(state<T>(cur), state<T>(1 - cur), kp, cell_filters ? d_decay : nullptr,
                                              cell_filters ? d_action : nullptr, mvar, d_epart, stream);
            });
            check_launch();
        }
        const bool summed = nw > 0;             // nw <= n_epart: sized for the shortest chunks
        if (summed) {
            launch_reduce_partials(d_epart, nw, d_partials, d_red, stream);
            check_launch();
        } else {
            do_maxwell<T>();
        }
        do_filters<T>();
        cur = 1 - cur;            // commit
        step += 1;
        if (summed) {
            energy_ready = true;
            energy_step = step;
        }
    }
    void mxz_steps(int L, bool energy_next) {
        fp32 ? do_mxz_steps<float>(L, energy_next) : do_mxz_steps<double>(L, energy_next);
    }
    // z sweep of both species then Maxwell, as the split kernels (v = 0) or
    // the fused kernel (v = 1..3); writes the next state
    template <typename T> void zm_launch(int v) {
        const double* dec = cell_filters ? d_decay : nullptr;
        const int* actp = cell_filters ? d_action : nullptr;
        if (v == 0) {
            fct_z<T>(0, nsp, zvar);
            launch_maxwell<T>(state<T>(cur), state<T>(1 - cur), kp, dec, actp, mvar, stream);
        } else {
            launch_fct_zmx<T>(state<T>(cur), stg<T>(), state<T>(1 - cur), kp, dec, actp, v, stream);
        }
    }
    template <typename T> void do_zm() {
        set_order(1);
        timed(K_FCT_ZM, [&] {
            launch_fct_zmx<T>(state<T>(cur), stg<T>(), state<T>(1 - cur), kp, cell_filters ? d_decay : nullptr,
                              cell_filters ? d_action : nullptr, zmvar, stream);
        });
        check_launch();
    }
    template <typename T> void do_maxwell() {
        set_order(2);
        timed(K_MAXWELL, [&] {
            if (emhd)
                launch_emhd_maxwell<T>(state<T>(cur), state<T>(1 - cur), kp, mvar, stream);
            else
                launch_maxwell<T>(state<T>(cur), state<T>(1 - cur), kp, cell_filters ? d_decay : nullptr,
                                  cell_filters ? d_action : nullptr, mvar, stream);
        });
        check_launch();
    }
    template <typename T> void do_modes(const ModeTable& t, const ArrayList<T>& l, int project) {
        if (t.nmodes == 0 || l.n == 0) return;
        if (t.is_float)
            launch_column_modes<T, float>(l, (const float*)t.c, (const float*)t.s, (const float*)t.w,
                                          t.nmodes, project, d_coef, kp, stream);
        else
            launch_column_modes<T, double>(l, (const double*)t.c, (const double*)t.s, (const double*)t.w,
                                           t.nmodes, project, d_coef, kp, stream);
    }
    template <typename T> void do_filters() {
        if (emhd) return;                 // the EMHD step has no filters
        const int k = 1 - cur;
        const khs_config& c = cfg;
        const bool skip_c1 = c.suppress_kz0 != 0;   // colour-1 EM is zeroed at the end anyway
        bool any = cell_filters || c.suppress_kz0 || c.kz_suppress_max >= 1 ||
                   c.kz_suppress_hi > c.kz_suppress_max + 1 || c.kz_keep_max > 0 || c.hyp_diff > 0.0f;
        if (!any) return;
        const bool bands = c.kz_suppress_max >= 1 || c.kz_suppress_hi > c.kz_suppress_max + 1;
        if (c.suppress_kz0 && !bands && c.kz_keep_max == 0 && c.hyp_diff <= 0.0f) {
            // z-mean only: charge cut, z-mean of the 12 colour-2/3 arrays
            // (kFilt15 without colour 1, charges last) and colour-1 zeroing in
            // one launch; the same filters as below, with the z-mean sums
            // in another order
            timed(K_FILTERS, [&] {
                int ids[15];
                int n = 0;
                for (int id : kFilt15)
                    if (id != S_BY1 && id != S_EX1 && id != S_EZ1) ids[n++] = id;
                ids[n++] = S_BY1; ids[n++] = S_EX1; ids[n++] = S_EZ1;
                launch_zmean_cut<T>(list<T>(k, ids, 15, false), 12, cell_filters ? 4 : 0, d_decay,
                                    cell_filters ? d_action : nullptr, kp, stream);
            });
            check_launch();
            return;
        }
        timed(K_FILTERS, [&] {
            // colour-2/3 fields were damped / cut in k_maxwell; the charge cut
            // here, and with suppress_kz0 the colour-1 EM zeroing (no later
            // filter touches colour 1 then), in one launch
            if (cell_filters || c.suppress_kz0) {
                static const int cut_c1[7] = {S_Q2A, S_Q3A, S_Q2B, S_Q3B, S_BY1, S_EX1, S_EZ1};
                const int ncut = cell_filters ? 4 : 0, nzero = c.suppress_kz0 ? 3 : 0;
                launch_cell_filters<T>(list<T>(k, cut_c1 + 4 - ncut, ncut + nzero, false), d_decay,
                                       cell_filters ? d_action : nullptr, kp, stream, nzero);
            }
            // With the MFMA keep projection (and no band filter between
            // them) the z-mean removal of the 15 field / charge arrays is
            // folded into it: those arrays come first in its list and are
            // synthesised without their kz = 0 term.  The two projections
            // commute, P_keep (1 - P_0) v = P_{1..K} v, and only rounding
            // differs from removing the mean first.  The bands use the
            // reference's FP32 twiddles, whose sums over z are not exactly
            // zero, so with bands the mean is removed before them.
            const bool fold_zmean = c.suppress_kz0 && c.kz_keep_max > 0 && d_keep_tab && !bands;
            if (c.suppress_kz0 && !fold_zmean) do_modes<T>(t_zmean, list<T>(k, kFilt15, 15, skip_c1), 0);
            // reference band filter: 15 field/charge arrays + (pz, px, n) x 2 species
            int ids21[21];
            std::memcpy(ids21, kFilt15, sizeof(kFilt15));
            std::memcpy(ids21 + 15, kFluid6, sizeof(kFluid6));
            if (c.kz_suppress_max >= 1) do_modes<T>(t_lo, list<T>(k, ids21, 21, skip_c1), 0);
            if (c.kz_suppress_hi > c.kz_suppress_max + 1) do_modes<T>(t_hi, list<T>(k, ids21, 21, skip_c1), 0);
            if (c.kz_keep_max > 0) {
                if (d_keep_tab) {
                    const ArrayList<T> keep = list<T>(k, ids21, 21, skip_c1);
                    const int nomean = fold_zmean ? keep.n - 6 : 0;     // all but the 6 fluid arrays
                    launch_keep_mfma<T>(keep, (const T*)d_keep_tab, (T*)d_coef, kp, stream, nomean);
                } else {
                    do_modes<T>(t_keep, list<T>(k, ids21, 21, skip_c1), 1);
                }
            }
            if (c.hyp_diff > 0.0f) launch_hyperdiffuse<T>(list<T>(k, kCol23, 12, false), (double)c.hyp_diff, kp, stream);
            // (colour-1 EM zeroing of suppress_kz0: with the charge cut above)
        });
        check_launch();
    }
    void x_stage(int sp0, int nsp) {
        autotune();
        fp32 ? do_x<float>(sp0, nsp) : do_x<double>(sp0, nsp);
    }
    void z_stage(int sp0, int nsp) { fp32 ? do_z<float>(sp0, nsp) : do_z<double>(sp0, nsp); }
    void xz_stage() {
        autotune();
        fp32 ? do_xz<float>() : do_xz<double>();
    }
    void maxwell_stage() { fp32 ? do_maxwell<float>() : do_maxwell<double>(); }
    void zm_stage() { fp32 ? do_zm<float>() : do_zm<double>(); }
    void filter_stage() { fp32 ? do_filters<float>() : do_filters<double>(); }

    // Energy reduction and its copy into the pinned h_red[0], without waiting
    // (also used inside a captured graph).
    void energy_async() {
        if (energy_ready && energy_step == step) {   // summed by the stride's last Maxwell
            energy_ready = false;
            HIPCHK(hipMemcpyAsync(h_red, d_red, sizeof(double), hipMemcpyDeviceToHost, stream));
            return;
        }
        timed(K_ENERGY, [&] {
            if (emhd) {
                if (fp32) launch_energy_em<float>(state<float>(cur), kp, d_partials, d_red, stream);
                else launch_energy_em<double>(state<double>(cur), kp, d_partials, d_red, stream);
            } else {
                if (fp32) launch_energy<float>(state<float>(cur), kp, d_partials, d_red, stream);
                else launch_energy<double>(state<double>(cur), kp, d_partials, d_red, stream);
            }
        });
        check_launch();
        HIPCHK(hipMemcpyAsync(h_red, d_red, sizeof(double), hipMemcpyDeviceToHost, stream));
    }
    double energy() {
        energy_async();
        HIPCHK(hipStreamSynchronize(stream));
        return h_red[0];
    }
};

Ctx::~Ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto g : sgraph)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto& p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : evpool) (void)hipEventDestroy(e);
    if (dmem) (void)hipFree(dmem);
    if (spec_mem) (void)hipFree(spec_mem);
    if (h_red) (void)hipHostFree(h_red);
    if (stream) (void)hipStreamDestroy(stream);
}

// --------------------------------------------------------------------------
static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

static float band_phase(int m, int z, int nz) {     // YM_Physics.cu:387
    float ph = 2.0f * (float)M_PI;
    ph = ph * (float)m;
    ph = ph * (float)z;
    ph = ph / (float)nz;
    return ph;
}

static std::unique_ptr<Ctx> make_ctx(const khs_config& in) {
    auto ctx = std::make_unique<Ctx>();
    ctx->cfg = in;
    config_derive(ctx->cfg);          // validates; idempotent on a derived config
    const khs_config& c = ctx->cfg;
    ctx->nx = c.nx; ctx->nz = c.nz;
    ctx->emhd = (c.model == KHS_MODEL_EMHD);
    ctx->nsp = ctx->emhd ? c.species : 2;
    ctx->N = (size_t)c.nx * c.nz;
    ctx->fp32 = (c.precision == KHS_FP32);
    ctx->esz = ctx->fp32 ? 4 : 8;
    ctx->stride = round_up(ctx->N * ctx->esz, 256);

    KParams& kp = ctx->kp;
    kp.nx = c.nx; kp.nz = c.nz;
    kp.periodic_x = c.periodic_x;
    kp.freeze_az1 = c.freeze_az1;
    kp.dt = c.dt; kp.dx = c.dx; kp.dz = c.dz;
    kp.dtdx = c.dt / c.dx;
    kp.dtdz = c.dt / c.dz;
    kp.c2 = c.c * c.c;
    kp.eps0 = c.eps_0;
    kp.rdx = 1.0 / c.dx;
    kp.rdz = 1.0 / c.dz;
    kp.reps0 = 1.0 / c.eps_0;
    kp.alpha = c.alpha;
    kp.warm_T = (double)c.warm_T;
    kp.q0 = c.charge1;
    kp.q1 = c.charge2;
    kp.nsp = ctx->nsp;

    int count;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
        throw Error(KHS_ERR_DEVICE, "no HIP device available");
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));

    // ---- mode tables (host) ----
    const int nz = c.nz;
    auto float_band = [&](int lo, int hi, std::vector<float>& hc, std::vector<float>& hs, std::vector<float>& hw) {
        int M = std::max(0, hi - lo + 1);
        hc.resize((size_t)M * nz); hs.resize((size_t)M * nz); hw.assign(M, 2.0f / (float)nz);
        for (int j = 0; j < M; ++j)
            for (int z = 0; z < nz; ++z) {
                float ph = band_phase(lo + j, z, nz);
                hc[(size_t)j * nz + z] = (float)std::cos((double)ph);
                hs[(size_t)j * nz + z] = (float)std::sin((double)ph);
            }
        return M;
    };
    std::vector<float> lo_c, lo_s, lo_w, hi_c, hi_s, hi_w;
    int nlo = (c.kz_suppress_max >= 1) ? float_band(1, c.kz_suppress_max, lo_c, lo_s, lo_w) : 0;
    int nhi = (c.kz_suppress_hi > c.kz_suppress_max + 1)
                  ? float_band(c.kz_suppress_max + 2, c.kz_suppress_hi, hi_c, hi_s, hi_w) : 0;
    std::vector<double> kp_c, kp_s, kp_w;
    int nkeep = 0;
    if (c.kz_keep_max > 0) {
        nkeep = c.kz_keep_max + 1;
        kp_c.resize((size_t)nkeep * nz); kp_s.resize((size_t)nkeep * nz); kp_w.resize(nkeep);
        for (int m = 0; m < nkeep; ++m) {
            kp_w[m] = (m == 0) ? 1.0 / nz : 2.0 / nz;
            for (int z = 0; z < nz; ++z) {
                double ph = 2.0 * M_PI * (double)((long)m * z % nz) / (double)nz;
                kp_c[(size_t)m * nz + z] = std::cos(ph);
                kp_s[(size_t)m * nz + z] = std::sin(ph);
            }
        }
    }
    std::vector<double> zm_c(nz, 1.0), zm_s(nz, 0.0), zm_w(1, 1.0 / (double)nz);

    // sponge / xi_cut per-column action (YM_Physics.cu:161-210)
    std::vector<double> decay(c.nx, 1.0);
    std::vector<int> action(c.nx, 0);
    for (int x = 0; x < c.nx; ++x) {
        double x_c = (c.nx * c.dx) * 0.5;
        double xi = std::fabs((x * c.dx - x_c) / c.eps);
        if (c.xi_sponge > 0.0f && xi > (double)c.xi_sponge) {
            double sigma = (double)c.sigma_sponge * (xi - (double)c.xi_sponge);
            decay[x] = std::exp(-sigma * c.dt);
            action[x] = 1;
        }
        if (c.xi_cut > 0.0f && xi > (double)c.xi_cut) action[x] = 2;
    }
    ctx->cell_filters = (c.xi_sponge > 0.0f) || (c.xi_cut > 0.0f);
    if (c.x_segment) ctx->xvar = c.x_segment == 64 ? 0 : c.x_segment == 128 ? 1 : 2;
    ctx->xvar += 3 * (c.x_rows == 1 ? 0 : c.x_rows == 16 ? 2 : 1) + 9 * (c.x_serial == 0 ? 0 : 1);
    if (c.model != KHS_MODEL_EMHD && c.x_lane_cells > 0) ctx->xvar = c.x_lane_cells == 4 ? 18 : 19;
    if (c.z_chunk) ctx->zvar = c.z_chunk == 16 ? 0 : c.z_chunk == 32 ? 1 : 2;
    if (c.z_prefetch == 1) ctx->zvar = ctx->zvar % 3 + 3;
    if (c.z_prefetch == 2) ctx->zvar = ctx->zvar % 3 + 12;
    if (c.z_groups == 1) ctx->zvar += 6;
    if (c.maxwell_kernel >= 0) ctx->mvar = c.maxwell_kernel;
    if (c.xz_variant >= 0) ctx->xzvar = c.xz_variant;
    if (c.model != KHS_MODEL_EMHD && c.zm_fused > 0) ctx->zmvar = c.zm_fused;

    // ---- one device allocation ----
    const bool share_az1 = c.freeze_az1 != 0;
    size_t off = 0;
    std::vector<std::pair<void**, size_t>> plan;
    auto reserve = [&](void** p, size_t bytes) { plan.push_back({p, off}); off += round_up(bytes, 256); };
    // The state and stage arrays are staggered by kArrayStagger bytes: a
    // kernel reads the same cell of some 24 arrays at once, and with array
    // sizes a multiple of a large power of two (16 MiB at 2048x1024 FP64)
    // those addresses agree in their low bits and fall on the same HBM
    // channels.  2048x1024 FP64: the Maxwell + next-sweep launch 225 -> 207
    // us with a 256-byte stagger (512, 1536, 2304, 4352: the same; 1024:
    // 233 us); FP32 136 -> 131 us.
    constexpr size_t kArrayStagger = 256;
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < S_COUNT; ++i) {
            if (k == 1 && i == S_AZ1 && share_az1) continue;
            reserve(&ctx->st[k][i], ctx->stride + kArrayStagger);
        }
    for (int i = 0; i < 2 * F_COUNT; ++i) reserve(&ctx->stage[i], ctx->stride + kArrayStagger);
    reserve((void**)&ctx->d_decay, c.nx * sizeof(double));
    reserve((void**)&ctx->d_action, c.nx * sizeof(int));
    auto reserve_tab = [&](ModeTable& t, int M, bool isf) {
        t.nmodes = M; t.is_float = isf;
        if (!M) return;
        size_t e = isf ? 4 : 8;
        reserve(&t.c, (size_t)M * nz * e);
        reserve(&t.s, (size_t)M * nz * e);
        reserve(&t.w, (size_t)M * e);
    };
    reserve_tab(ctx->t_lo, nlo, true);
    reserve_tab(ctx->t_hi, nhi, true);
    reserve_tab(ctx->t_keep, nkeep, false);
    reserve_tab(ctx->t_zmean, c.suppress_kz0 ? 1 : 0, false);
    int maxm = std::max(std::max(nlo, nhi), std::max(nkeep, c.suppress_kz0 ? 1 : 0));
    const bool keep_mfma = c.mfma_filter && c.kz_keep_max > 0 && keep_mfma_supported(c.nx, c.nz, c.kz_keep_max);
    // coefficient scratch: 21 arrays x 2 x modes doubles (column kernels) or
    // 21 arrays x 17 rows of the run precision (MFMA keep filter)
    size_t coef_bytes = (size_t)21 * 2 * maxm * c.nx * sizeof(double);
    if (keep_mfma) coef_bytes = std::max(coef_bytes, (size_t)21 * 17 * c.nx * ctx->esz);
    if (coef_bytes) reserve((void**)&ctx->d_coef, coef_bytes);
    if (keep_mfma) reserve((void**)&ctx->d_keep_tab, (size_t)c.nz * 16 * ctx->esz);
    reserve((void**)&ctx->d_partials, (size_t)kRedBlocks * 8 * 2 * sizeof(double));
    // the Maxwell march's per-wavefront energy sums: 63 columns x 2 rows at least
    if (!ctx->emhd) {
        ctx->n_epart = ((c.nx + 62) / 63) * ((c.nz + 1) / 2);
        reserve((void**)&ctx->d_epart, (size_t)ctx->n_epart * 2 * sizeof(double));
    }
    reserve((void**)&ctx->d_red, 16 * sizeof(double));
    reserve((void**)&ctx->d_derived, ctx->N * sizeof(double));

    HIPCHK(hipMalloc((void**)&ctx->dmem, off));
    for (auto& pr : plan) *pr.first = ctx->dmem + pr.second;
    if (share_az1) ctx->st[1][S_AZ1] = ctx->st[0][S_AZ1];
    HIPCHK(hipMemsetAsync(ctx->dmem, 0, off, ctx->stream));
    HIPCHK(hipHostMalloc((void**)&ctx->h_red, 16 * sizeof(double), hipHostMallocDefault));

    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    };
    up(ctx->d_decay, decay.data(), decay.size() * sizeof(double));
    up(ctx->d_action, action.data(), action.size() * sizeof(int));
    if (nlo) { up(ctx->t_lo.c, lo_c.data(), lo_c.size() * 4); up(ctx->t_lo.s, lo_s.data(), lo_s.size() * 4); up(ctx->t_lo.w, lo_w.data(), lo_w.size() * 4); }
    if (nhi) { up(ctx->t_hi.c, hi_c.data(), hi_c.size() * 4); up(ctx->t_hi.s, hi_s.data(), hi_s.size() * 4); up(ctx->t_hi.w, hi_w.data(), hi_w.size() * 4); }
    if (nkeep) { up(ctx->t_keep.c, kp_c.data(), kp_c.size() * 8); up(ctx->t_keep.s, kp_s.data(), kp_s.size() * 8); up(ctx->t_keep.w, kp_w.data(), kp_w.size() * 8); }
    std::vector<float> kt;            // live until the synchronisation below
    std::vector<double> kt64;
    if (keep_mfma && ctx->fp32) {
        kt.resize((size_t)c.nz * 16);
        fill_keep_mfma_table<float>(c.nz, c.kz_keep_max, kt.data());
        up(ctx->d_keep_tab, kt.data(), kt.size() * sizeof(float));
    } else if (keep_mfma) {
        kt64.resize((size_t)c.nz * 16);
        fill_keep_mfma_table<double>(c.nz, c.kz_keep_max, kt64.data());
        up(ctx->d_keep_tab, kt64.data(), kt64.size() * sizeof(double));
    }
    if (c.suppress_kz0) { up(ctx->t_zmean.c, zm_c.data(), nz * 8); up(ctx->t_zmean.s, zm_s.data(), nz * 8); up(ctx->t_zmean.w, zm_w.data(), 8); }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ctx;
}

static void upload_state(Ctx& ctx, int k, int id, const double* h) {
    void* dst = ctx.st[k][id];
    if (!ctx.fp32) {
        HIPCHK(hipMemcpyAsync(dst, h, ctx.N * 8, hipMemcpyHostToDevice, ctx.stream));
        HIPCHK(hipStreamSynchronize(ctx.stream));
    } else {
        std::vector<float> tmp(ctx.N);
        for (size_t i = 0; i < ctx.N; ++i) tmp[i] = (float)h[i];
        HIPCHK(hipMemcpyAsync(dst, tmp.data(), ctx.N * 4, hipMemcpyHostToDevice, ctx.stream));
        HIPCHK(hipStreamSynchronize(ctx.stream));
    }
}

static void download(Ctx& ctx, const void* src, bool isf, double* h) {
    if (!isf) {
        HIPCHK(hipMemcpyAsync(h, src, ctx.N * 8, hipMemcpyDeviceToHost, ctx.stream));
        HIPCHK(hipStreamSynchronize(ctx.stream));
    } else {
        std::vector<float> tmp(ctx.N);
        HIPCHK(hipMemcpyAsync(tmp.data(), src, ctx.N * 4, hipMemcpyDeviceToHost, ctx.stream));
        HIPCHK(hipStreamSynchronize(ctx.stream));
        for (size_t i = 0; i < ctx.N; ++i) h[i] = (double)tmp[i];
    }
}

// State slot of an array id of the context's model, or -1.
//   YM   : ids 0..23 are the slots.
//   EMHD : ids 64..72 fields -> slots 0..8; 76..79 / 82..85 fluid n px py pz
//          of species 0 / 1 -> slots 12..15 / 18..21.
static int state_slot(const Ctx& c, int id) {
    if (!c.emhd) return (id >= 0 && id < KHS_N_STATE) ? id : -1;
    if (id >= KHS_EM_EX && id <= KHS_EM_AZ) return id - 64;
    if (id >= KHS_EM_N && id <= KHS_EM_PZ) return id - 64;
    if (c.nsp > 1 && id >= KHS_EM_N2 && id <= KHS_EM_PZ2) return id - 64;
    return -1;
}

// Stage slot of a fluid state slot (species sp, variable v -> sp * NV + v).
static int stage_slot(const Ctx& c, int slot) {
    if (!c.emhd) return slot >= S_NA ? slot - S_NA : -1;
    if (slot < 12) return -1;
    const int sp = (slot - 12) / 6, v = (slot - 12) % 6;
    return (v < EV_COUNT) ? sp * EV_COUNT + v : -1;
}

static void require_idle(const Ctx& c, const char* what) {
    if (c.staged_x || c.staged_z || c.maxwell_done || c.filters_done)
        throw Error(KHS_ERR_STATE, std::string(what) + " called in the middle of a staged step (call khs_commit first)");
}

static void commit(Ctx& c) {
    c.cur = 1 - c.cur;
    c.staged_x = c.staged_z = 0;
    c.maxwell_done = c.filters_done = false;
    c.step += 1;
}

static void full_step(Ctx& c);
static void run_steps(Ctx& c, int L, bool energy_next = false);

// Capture `stride` full steps and the energy reduction (into the pinned
// h_red[0]) as one HIP graph, starting from ping-pong buffer c.cur, into
// c.sgraph[c.cur].  An even stride starts and ends on the same buffer; an odd
// one alternates, so it is captured once from each buffer (the tile-walk
// direction of each step, set by the step's parity, is the same on every
// replay because buffer and parity alternate together).
static void capture_stride(Ctx& c, int stride) {
    hipGraphExec_t& slot = c.sgraph[c.cur];
    if (slot) (void)hipGraphExecDestroy(slot);
    slot = nullptr;
    c.autotune();                               // tuning launches and waits: before capture
    HIPCHK(hipStreamSynchronize(c.stream));
    const int cur0 = c.cur;
    const long step0 = c.step;
    std::lock_guard<std::mutex> lk(capture_mutex());
    HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
    try {
        run_steps(c, stride, true);
        c.energy_async();
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c.stream, &g);       // leave the stream usable
        if (g) (void)hipGraphDestroy(g);
        c.cur = cur0;
        c.step = step0;
        c.energy_ready = false;
        throw;
    }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(c.stream, &g));
    c.cur = cur0;                               // the graph does the stepping
    c.step = step0;
    hipError_t e = hipGraphInstantiate(&slot, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        slot = nullptr;
        throw Error(KHS_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
}

// Whole energy strides run as one captured graph (no per-kernel launch cost
// on the host); kernel timing needs the eager path.
static bool stride_graphs(const Ctx& c) {
    return c.cfg.step_graphs && !c.timing && c.cfg.energy_stride >= 2;
}

static void full_step(Ctx& c) {
    c.autotune();                               // decides split or fused sweeps on the first step
    if (c.fused) {
        c.xz_stage();
        c.maxwell_stage();
    } else {
        c.x_stage(0, c.nsp);                    // autotunes on the first call
        if (c.zmvar > 0) {
            c.zm_stage();
        } else {
            c.z_stage(0, c.nsp);
            c.maxwell_stage();
        }
    }
    c.filter_stage();
    commit(c);
}

// L consecutive full steps without an energy check between them;
// energy_next: the caller reads the energy of the resulting state next
// (energy_async), which the last step may then sum as it writes the state.
static void run_steps(Ctx& c, int L, bool energy_next) {
    c.autotune();
    c.energy_ready = false;
    if (c.mxz && L >= 2) {
        c.mxz_steps(L, energy_next);
        return;
    }
    for (int s = 0; s < L; ++s) full_step(c);
}

static void write_energy_row(Ctx& c, long step, double E) {
    if (!c.ecsv.is_open()) return;
    double ratio = E / c.E0;
    double t_now = step * c.cfg.dt;
    c.ecsv << step << ',' << t_now << ',' << E << ',' << ratio << '\n';
    c.ecsv.flush();
}

// --------------------------------------------------------------------------
// Batch of independent runs (parameter sweeps).  The runs are spread over
// `lanes` streams; K steps of every active run, followed by each run's energy
// reduction and its copy to pinned host memory, are captured once into a HIP
// graph and replayed, so a K-step segment of the whole sweep is one graph
// launch.  Halting (NaN/Inf or E/E0 above the run's threshold) is checked on
// the host after each segment; a halted run leaves the graph, which is then
// captured again.  K is even so that every replay starts from the same
// ping-pong buffers.
// --------------------------------------------------------------------------
struct Batch {
    std::vector<std::unique_ptr<Ctx>> runs;
    // packed launches: descriptors of both step parities and tile offsets
    char* pk_mem = nullptr;
    size_t pk_bytes = 0;
    int pk_shape = -1;                        // packed launch shape, timed on the first capture
    std::vector<hipStream_t> lanes;
    hipStream_t origin = nullptr;
    std::vector<hipEvent_t> ev;              // fork event + one join event per lane
    int seg = 50;                             // steps per graph segment (even)
    std::vector<int> active;                  // 1 while a run has not halted
    std::vector<long> cap;                    // a run stops (not halted) at this step
    hipGraphExec_t exec = nullptr;
    int exec_k = 0;
    std::vector<int> exec_set;                // runs captured in exec
    std::vector<int> exec_cur;                // their ping-pong buffers when captured
    bool exec_packed = false;                 // exec holds packed launches
    std::vector<int> cur_of(const std::vector<int>& set) const {
        std::vector<int> v;
        for (int i : set) v.push_back(runs[i]->cur);
        return v;
    }
    ~Batch();
    void release_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        exec_k = 0;
        exec_set.clear();
        exec_cur.clear();
        exec_packed = false;
    }
    void capture(int K, const std::vector<int>& set) {
        release_graph();
        if (set.empty()) return;
        for (int i : set) runs[i]->autotune();          // tuning launches and waits: before capture
        for (int i : set) HIPCHK(hipStreamSynchronize(runs[i]->stream));
        std::lock_guard<std::mutex> lk(capture_mutex());
        HIPCHK(hipStreamBeginCapture(origin, hipStreamCaptureModeRelaxed));
        HIPCHK(hipEventRecord(ev[0], origin));
        for (auto l : lanes) HIPCHK(hipStreamWaitEvent(l, ev[0], 0));
        // runs dealt to lanes by longest processing time first (cells per
        // step), so that the lanes' chains of kernels end together
        std::vector<int> order(set.size()), lane_of(set.size());
        for (size_t k = 0; k < set.size(); ++k) order[k] = (int)k;
        auto cost = [&](int k) { return (double)runs[set[k]]->nx * (double)runs[set[k]]->nz; };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
        std::vector<double> load(lanes.size(), 0.0);
        for (int k : order) {
            const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            lane_of[k] = l;
            load[l] += cost(k);
        }
        for (int k : order) {
            Ctx& c = *runs[set[k]];
            hipStream_t own = c.stream;
            c.stream = lanes[lane_of[k]];
            const int cur0 = c.cur;
            const long step0 = c.step;
            run_steps(c, K, true);
            c.energy_async();
            c.cur = cur0;                                   // the graph does the stepping
            c.step = step0;
            c.stream = own;
        }
        for (size_t l = 0; l < lanes.size(); ++l) {
            HIPCHK(hipEventRecord(ev[1 + l], lanes[l]));
            HIPCHK(hipStreamWaitEvent(origin, ev[1 + l], 0));
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamEndCapture(origin, &g));
        hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            exec = nullptr;
            throw Error(KHS_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        exec_k = K;
        exec_set = set;
        exec_cur = cur_of(set);
    }

    // Runs that packed launches can step: the YM model, one precision, and a
    // filter stage of at most the charge cut / sponge and the z-mean removal
    // (the kernels of the packed filter launch).
    bool packable(const std::vector<int>& set) const {
        if (set.empty()) return false;
        const bool f32 = runs[set[0]]->fp32;
        for (int i : set) {
            const Ctx& c = *runs[i];
            const khs_config& k = c.cfg;
            if (c.emhd || c.fp32 != f32) return false;
            if (k.kz_suppress_max >= 1 || k.kz_suppress_hi > k.kz_suppress_max + 1 || k.kz_keep_max > 0 ||
                k.hyp_diff > 0.0f)
                return false;
        }
        return true;
    }
    template <typename T> void capture_packed_t(int K, const std::vector<int>& set) {
        const int n = (int)set.size();
        std::vector<PackedRun<T>> d(2 * (size_t)n);
        std::vector<int> fb(n + 1, 0);
        for (int j = 0; j < n; ++j) {
            Ctx& c = *runs[set[j]];
            const int nstrip = (c.nx + 63) / 64;
            for (int par = 0; par < 2; ++par) {
                PackedRun<T>& r = d[(size_t)par * n + j];
                const int k = c.cur ^ par;                  // the buffer step parity `par` starts from
                r.S = c.state<T>(k);
                r.O = c.state<T>(1 - k);
                r.p = c.kp;
                r.p.rev = 0;
                r.decay = c.cell_filters ? c.d_decay : nullptr;
                r.action = c.cell_filters ? c.d_action : nullptr;
                r.ngx = (c.nx + 57) / 58;
                r.nwx = (c.nx + 62) / 63;
                r.nstrip = nstrip;
                r.fmode = 0;
                r.nmean = r.nact = 0;
                r.filt = ArrayList<T>{};
                if (c.cfg.suppress_kz0) {
                    // as the one-launch z-mean filter of do_filters
                    int ids[15];
                    int m = 0;
                    for (int id : kFilt15)
                        if (id != S_BY1 && id != S_EX1 && id != S_EZ1) ids[m++] = id;
                    ids[m++] = S_BY1; ids[m++] = S_EX1; ids[m++] = S_EZ1;
                    r.filt = c.list<T>(1 - k, ids, 15, false);
                    r.fmode = 2;
                    r.nmean = 12;
                    r.nact = c.cell_filters ? 4 : 0;
                } else if (c.cell_filters) {
                    r.filt = c.list<T>(1 - k, kQ23, 4, false);
                    r.fmode = 1;
                }
            }
            fb[j + 1] = fb[j] + (d[j].fmode ? nstrip * d[j].filt.n : 0);
        }
        // tile offsets of the sweep and the Maxwell march for a launch shape
        auto offsets = [&](int shape, std::vector<int>& ix) {
            const int xc = packed_xz_chunk(shape), mc = packed_mx_chunk(shape);
            ix.assign(3 * (size_t)(n + 1), 0);
            for (int j = 0; j < n; ++j) {
                const Ctx& c = *runs[set[j]];
                ix[j + 1] = ix[j] + 2 * d[j].ngx * ((c.nz + xc - 1) / xc);
                ix[(n + 1) + j + 1] = ix[(n + 1) + j] + d[j].nwx * ((c.nz + mc - 1) / mc);
            }
            std::copy(fb.begin(), fb.end(), ix.begin() + 2 * (n + 1));
        };
        const size_t dbytes = round_up(d.size() * sizeof(PackedRun<T>), 256);
        const size_t ibytes = round_up(3 * (size_t)(n + 1) * sizeof(int), 256);
        if (pk_bytes < dbytes + ibytes) {
            if (pk_mem) (void)hipFree(pk_mem);
            pk_mem = nullptr;
            HIPCHK(hipMalloc((void**)&pk_mem, dbytes + ibytes));
            pk_bytes = dbytes + ibytes;
        }
        HIPCHK(hipMemcpy(pk_mem, d.data(), d.size() * sizeof(PackedRun<T>), hipMemcpyHostToDevice));
        const PackedRun<T>* dev = reinterpret_cast<const PackedRun<T>*>(pk_mem);
        const int* ib = reinterpret_cast<const int*>(pk_mem + dbytes);
        std::vector<int> ix;
        auto step = [&](int st, int shape) {
            launch_packed_step<T>(dev + (size_t)(st & 1) * n, ib, ib + (n + 1), ib + 2 * (n + 1), n, ix[n],
                                  ix[(n + 1) + n], ix[2 * (n + 1) + n], shape, origin);
        };
        if (pk_shape < 0 && runs[set[0]]->cfg.packed_shape >= 0) pk_shape = runs[set[0]]->cfg.packed_shape;
        if (pk_shape < 0) {
            // time every launch shape on this batch (a step from the current
            // state into the next-state buffers, which the first real step
            // overwrites) and keep the fastest for the batch's lifetime
            hipEvent_t ea, eb;
            HIPCHK(hipEventCreate(&ea));
            HIPCHK(hipEventCreate(&eb));
            float best = 1e30f;
            for (int shape = 0; shape < kPackedShapes; ++shape) {
                offsets(shape, ix);
                HIPCHK(hipMemcpy(pk_mem + dbytes, ix.data(), ix.size() * sizeof(int), hipMemcpyHostToDevice));
                step(0, shape);
                HIPCHK(hipEventRecord(ea, origin));
                for (int r = 0; r < 3; ++r) step(0, shape);
                HIPCHK(hipEventRecord(eb, origin));
                HIPCHK(hipEventSynchronize(eb));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, ea, eb));
                if (ms < best) { best = ms; pk_shape = shape; }
            }
            (void)hipEventDestroy(ea);
            (void)hipEventDestroy(eb);
        }
        offsets(pk_shape, ix);
        HIPCHK(hipMemcpy(pk_mem + dbytes, ix.data(), ix.size() * sizeof(int), hipMemcpyHostToDevice));

        std::lock_guard<std::mutex> lk(capture_mutex());
        HIPCHK(hipStreamBeginCapture(origin, hipStreamCaptureModeRelaxed));
        for (int st = 0; st < K; ++st) step(st, pk_shape);
        for (int i : set) {                         // each run's energy into its pinned h_red[0]
            Ctx& c = *runs[i];
            hipStream_t own = c.stream;
            c.stream = origin;
            c.energy_async();
            c.stream = own;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamEndCapture(origin, &g));
        hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            exec = nullptr;
            throw Error(KHS_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    }
    void capture_packed(int K, const std::vector<int>& set) {
        release_graph();
        if (set.empty()) return;
        for (int i : set) HIPCHK(hipStreamSynchronize(runs[i]->stream));
        HIPCHK(hipStreamSynchronize(origin));
        if (runs[set[0]]->fp32) capture_packed_t<float>(K, set);
        else capture_packed_t<double>(K, set);
        exec_k = K;
        exec_set = set;
        exec_cur = cur_of(set);
        exec_packed = true;
    }
};

Batch::~Batch() {
    release_graph();
    if (pk_mem) {
        if (origin) (void)hipStreamSynchronize(origin);
        (void)hipFree(pk_mem);
    }
    if (origin) (void)hipStreamSynchronize(origin);
    for (auto e : ev) (void)hipEventDestroy(e);
    for (auto l : lanes) (void)hipStreamDestroy(l);
    if (origin) (void)hipStreamDestroy(origin);
    runs.clear();
}

// --------------------------------------------------------------------------
// Device scratch of the allocation-free energy entry point: the partial sums
// of the energy reduction (one compensated pair per reduction block, at most
// kRedBlocks blocks whatever the grid), 16 KiB per GPU, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (in host
// code the name is a host-side stand-in, not a device address): the device
// address is looked up once per device with hipGetSymbolAddress, with that
// device current, and used from then on.
// --------------------------------------------------------------------------
constexpr size_t kScratchDoubles = (size_t)kRedBlocks * 2;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double* g_scratch_dev[kMaxDevices] = {};

static double* scratch_of_current_device() {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(KHS_ERR_DEVICE, "device ordinal out of range");
    static std::mutex m;
    std::lock_guard<std::mutex> lk(m);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(KHS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    return g_scratch_dev[dev];
}

}  // namespace khs

// ==========================================================================
// C ABI
// ==========================================================================
using khs::Ctx;
using khs::Error;

struct khs_ctx {
    std::unique_ptr<Ctx> impl;
};

static thread_local std::string g_last_error;

template <typename F>
static int guard(F&& f) {
    try {
        f();
        g_last_error.clear();
        return KHS_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return KHS_ERR_ARG;
    }
}

#define CTX_OR_FAIL(ctx)                                                 \
    if (!(ctx) || !(ctx)->impl) { g_last_error = "null context"; return KHS_ERR_ARG; }

extern "C" {

const char* khs_last_error(void) { return g_last_error.c_str(); }
const char* khs_version(void) { return "khs 0.1 (gfx950)"; }
size_t khs_config_sizeof(void) { return sizeof(khs_config); }

int khs_config_defaults(khs_config* cfg) {
    if (!cfg) { g_last_error = "null config"; return KHS_ERR_ARG; }
    return guard([&] { khs::config_defaults(*cfg); khs::config_derive(*cfg); });
}

int khs_config_set(khs_config* cfg, const char* key, const char* value) {
    if (!cfg || !key || !value) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    return guard([&] {
        if (!khs::config_set(*cfg, key, value))
            throw Error(KHS_ERR_CONFIG, std::string("unknown key '") + key + "'");
    });
}

int khs_load_ini(const char* path, khs_config* cfg) {
    if (!path || !cfg) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    return guard([&] {
        khs::config_defaults(*cfg);
        std::string warn;
        khs::config_load(path, *cfg, &warn);
        if (!warn.empty()) fputs(warn.c_str(), stderr);
    });
}

int khs_derive(khs_config* cfg) {
    if (!cfg) { g_last_error = "null config"; return KHS_ERR_ARG; }
    return guard([&] { khs::config_derive(*cfg); });
}

int khs_output_dir(const khs_config* cfg, char* buf, size_t buflen) {
    if (!cfg || !buf) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    return guard([&] {
        std::string s = khs::output_dir_name(*cfg);
        if (s.size() + 1 > buflen) throw Error(KHS_ERR_ARG, "buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

int khs_create(const khs_config* cfg, khs_ctx** out) {
    if (!cfg || !out) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    *out = nullptr;
    return guard([&] {
        auto h = new khs_ctx;
        try { h->impl = khs::make_ctx(*cfg); } catch (...) { delete h; throw; }
        *out = h;
    });
}

int khs_destroy(khs_ctx* ctx) {
    if (!ctx) return KHS_OK;
    int rc = guard([&] { ctx->impl.reset(); });
    delete ctx;
    return rc;
}

int khs_init(khs_ctx* ctx, const char* seed_path) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        std::string sp = seed_path ? std::string(seed_path) : std::string(c.cfg.seed_profile_file);
        khs::SeedProfiles seed;
        if (!sp.empty()) seed = khs::read_seed_file(sp, c.nx);
        std::vector<std::vector<double>> host;
        if (c.emhd) khs::host_initial_state_em(c.cfg, host);
        else khs::host_initial_state(c.cfg, seed, host);
        c.cur = 0;
        c.staged_x = c.staged_z = 0;
        c.maxwell_done = c.filters_done = false;
        for (int i = 0; i < khs::S_COUNT; ++i) khs::upload_state(c, 0, i, host[i].data());
        c.step = 0;
        c.E0 = c.energy();
    });
}

int khs_array_id(const char* name) {
    if (!name) return -1;
    static const char* names[] = {"Ex1", "Ex2", "Ex3", "Ez1", "Ez2", "Ez3", "By1", "By2", "By3",
                                  "Az1", "Az2", "Az3", "nA", "PxA", "PzA", "Q1A", "Q2A", "Q3A",
                                  "nB", "PxB", "PzB", "Q1B", "Q2B", "Q3B"};
    static const char* alt[] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                nullptr, nullptr, nullptr, nullptr, "pxA", "pzA", nullptr, nullptr, nullptr,
                                nullptr, "pxB", "pzB", nullptr, nullptr, nullptr};
    for (int i = 0; i < KHS_N_STATE; ++i) {
        if (std::strcmp(name, names[i]) == 0) return i;
        if (alt[i] && std::strcmp(name, alt[i]) == 0) return i;
    }
    static const char* dnames[] = {"Jx1", "Jx2", "Jx3", "Jz1", "Jz2", "Jz3", "uxA", "uzA", "uxB", "uzB",
                                   "srcpxA", "srcpzA", "srcpxB", "srcpzB", "sQ1A", "sQ2A", "sQ3A",
                                   "sQ1B", "sQ2B", "sQ3B", "energy_density"};
    for (int i = 0; i < (int)(sizeof(dnames) / sizeof(dnames[0])); ++i)
        if (std::strcmp(name, dnames[i]) == 0) return KHS_JX1 + i;
    // EMHD names (a name shared with the YM set resolves to the YM id above;
    // the EMHD energy density is "em_energy_density")
    static const char* enames[] = {"Ex", "Ey", "Ez", "Bx", "By", "Bz", "Ax", "Ay", "Az"};
    for (int i = 0; i < 9; ++i)
        if (std::strcmp(name, enames[i]) == 0) return KHS_EM_EX + i;
    static const char* fnames[] = {"n", "px", "py", "pz", "n2", "px2", "py2", "pz2"};
    for (int i = 0; i < 8; ++i)
        if (std::strcmp(name, fnames[i]) == 0) return (i < 4 ? KHS_EM_N : KHS_EM_N2) + (i & 3);
    static const char* ednames[] = {"Jx", "Jy", "Jz", "ux", "uz", "ux2", "uz2", "sx", "sy", "sz",
                                    "sx2", "sy2", "sz2", "em_energy_density"};
    for (int i = 0; i < 14; ++i)
        if (std::strcmp(name, ednames[i]) == 0) return KHS_EM_JX + i;
    return -1;
}

int khs_get_array(khs_ctx* ctx, int id, double* host, size_t n) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (!host || n != c.N) throw Error(KHS_ERR_ARG, "host buffer must hold nx*nz doubles");
        const int slot = khs::state_slot(c, id);
        if (slot >= 0) {
            khs::download(c, c.st[c.cur][slot], c.fp32, host);
        } else if (c.emhd && id >= KHS_EM_JX && id <= KHS_EM_ENERGY_DENSITY) {
            const int w = id - KHS_EM_JX;
            if (c.nsp < 2 && (w == 5 || w == 6 || (w >= 10 && w <= 12)))
                throw Error(KHS_ERR_ARG, "array id " + std::to_string(id) + " needs species = 2");
            if (c.fp32) khs::launch_derived_em<float>(c.state<float>(c.cur), w, c.kp, c.d_derived, c.stream);
            else khs::launch_derived_em<double>(c.state<double>(c.cur), w, c.kp, c.d_derived, c.stream);
            c.check_launch();
            khs::download(c, c.d_derived, false, host);
        } else if (!c.emhd && id >= KHS_JX1 && id < KHS_N_ARRAY_IDS) {
            if (c.fp32) khs::launch_derived<float>(c.state<float>(c.cur), id, c.kp, c.d_derived, c.stream);
            else khs::launch_derived<double>(c.state<double>(c.cur), id, c.kp, c.d_derived, c.stream);
            c.check_launch();
            khs::download(c, c.d_derived, false, host);
        } else {
            throw Error(KHS_ERR_ARG, "unknown array id " + std::to_string(id));
        }
    });
}

int khs_get_array_buf(khs_ctx* ctx, int buf, int id, double* host, size_t n) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (!host || n != c.N) throw Error(KHS_ERR_ARG, "host buffer must hold nx*nz doubles");
        const int slot = khs::state_slot(c, id);
        if (slot < 0) throw Error(KHS_ERR_ARG, "state ids of the model only");
        const void* src = nullptr;
        if (buf == KHS_BUF_CURRENT) src = c.st[c.cur][slot];
        else if (buf == KHS_BUF_NEXT) src = c.st[1 - c.cur][slot];
        else if (buf == KHS_BUF_STAGE) {
            const int g = khs::stage_slot(c, slot);
            if (g < 0) throw Error(KHS_ERR_ARG, "the stage holds fluid arrays only");
            src = c.stage[g];
        } else throw Error(KHS_ERR_ARG, "unknown buffer");
        khs::download(c, src, c.fp32, host);
    });
}

int khs_set_array(khs_ctx* ctx, int id, const double* host, size_t n) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (!host || n != c.N) throw Error(KHS_ERR_ARG, "host buffer must hold nx*nz doubles");
        const int slot = khs::state_slot(c, id);
        if (slot < 0) throw Error(KHS_ERR_ARG, "only state arrays of the model can be set");
        khs::require_idle(c, "khs_set_array");
        khs::upload_state(c, c.cur, slot, host);
    });
}

int khs_fct_x_sweep(khs_ctx* ctx, int species) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (species < 0 || species >= c.nsp) throw Error(KHS_ERR_ARG, "species out of range");
        if (c.staged_z || c.maxwell_done) throw Error(KHS_ERR_STATE, "x-sweep after z-sweep: commit the step first");
        c.x_stage(species, 1);
        c.staged_x |= (1 << species);
    });
}

int khs_fct_z_sweep(khs_ctx* ctx, int species) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (species < 0 || species >= c.nsp) throw Error(KHS_ERR_ARG, "species out of range");
        if (!(c.staged_x & (1 << species))) throw Error(KHS_ERR_STATE, "z-sweep before x-sweep of this species");
        c.z_stage(species, 1);
        c.staged_z |= (1 << species);
    });
}

int khs_fluid_velocities(khs_ctx* ctx, int species, double* ux, double* uz, size_t n) {
    CTX_OR_FAIL(ctx);
    Ctx& c = *ctx->impl;
    if (species < 0 || species >= c.nsp) { g_last_error = "species out of range"; return KHS_ERR_ARG; }
    const int idx = c.emhd ? (species ? KHS_EM_UX2 : KHS_EM_UX) : (species ? KHS_UXB : KHS_UXA);
    const int idz = c.emhd ? (species ? KHS_EM_UZ2 : KHS_EM_UZ) : (species ? KHS_UZB : KHS_UZA);
    const int rc = khs_get_array(ctx, idx, ux, n);
    return rc != KHS_OK ? rc : khs_get_array(ctx, idz, uz, n);
}

int khs_sources(khs_ctx* ctx, int species, double* src_px, double* src_pz, double* sQ, size_t n) {
    CTX_OR_FAIL(ctx);
    Ctx& c = *ctx->impl;
    if (species < 0 || species >= c.nsp) { g_last_error = "species out of range"; return KHS_ERR_ARG; }
    if (!sQ) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    int ids[5];
    int nq;
    if (c.emhd) {
        const int o = species ? 3 : 0;
        ids[0] = KHS_EM_SX + o; ids[1] = KHS_EM_SZ + o; ids[2] = KHS_EM_SY + o;
        nq = 1;
    } else {
        ids[0] = species ? KHS_SRCPXB : KHS_SRCPXA;
        ids[1] = species ? KHS_SRCPZB : KHS_SRCPZA;
        for (int k = 0; k < 3; ++k) ids[2 + k] = (species ? KHS_SQ1B : KHS_SQ1A) + k;
        nq = 3;
    }
    int rc = khs_get_array(ctx, ids[0], src_px, n);
    if (rc == KHS_OK) rc = khs_get_array(ctx, ids[1], src_pz, n);
    for (int k = 0; k < nq && rc == KHS_OK; ++k) rc = khs_get_array(ctx, ids[2 + k], sQ + (size_t)k * n, n);
    return rc;
}

int khs_enforce_boundaries(khs_ctx* ctx) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        khs::require_idle(c, "khs_enforce_boundaries");
        if (c.fp32) khs::launch_enforce_boundaries<float>(c.state<float>(c.cur), c.kp, c.emhd ? 1 : 0, c.stream);
        else khs::launch_enforce_boundaries<double>(c.state<double>(c.cur), c.kp, c.emhd ? 1 : 0, c.stream);
        c.check_launch();
        HIPCHK(hipStreamSynchronize(c.stream));
    });
}

int khs_maxwell(khs_ctx* ctx) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (c.staged_z != (1 << c.nsp) - 1) throw Error(KHS_ERR_STATE, "khs_maxwell needs every species z-swept");
        c.maxwell_stage();
        c.maxwell_done = true;
    });
}

int khs_filters(khs_ctx* ctx) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (!c.maxwell_done) throw Error(KHS_ERR_STATE, "khs_filters needs khs_maxwell first");
        c.filter_stage();
        c.filters_done = true;
    });
}

int khs_commit(khs_ctx* ctx) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (!c.maxwell_done) throw Error(KHS_ERR_STATE, "khs_commit needs khs_maxwell first");
        if (!c.filters_done) c.filter_stage();
        khs::commit(c);
    });
}

int khs_step(khs_ctx* ctx, int nsteps, int* halted) {
    CTX_OR_FAIL(ctx);
    if (halted) *halted = 0;
    return guard([&] {
        Ctx& c = *ctx->impl;
        khs::require_idle(c, "khs_step");
        const int stride = c.cfg.energy_stride;
        const bool graphs = khs::stride_graphs(c);
        for (int s = 0; s < nsteps;) {
            if (graphs && c.step % stride == 0 && nsteps - s >= stride) {
                if (!c.sgraph[c.cur]) khs::capture_stride(c, stride);
                HIPCHK(hipGraphLaunch(c.sgraph[c.cur], c.stream));
                HIPCHK(hipStreamSynchronize(c.stream));
                c.step += stride;
                c.cur ^= (stride & 1);
                s += stride;
                const double E = c.h_red[0];
                c.last_check_E = E;
                c.last_check_step = c.step;
                khs::write_energy_row(c, c.step, E);
                if (std::isnan(E) || std::isinf(E) || E / c.E0 > c.cfg.energy_thresh) {
                    if (halted) *halted = 1;
                    break;
                }
                continue;
            }
            // the steps up to the next energy check (or the end) as one segment
            long L = nsteps - s;
            if (stride > 0) L = std::min<long>(L, stride - c.step % stride);
            khs::run_steps(c, (int)L, stride > 0 && (c.step + L) % stride == 0);
            s += (int)L;
            if (stride > 0 && c.step % stride == 0) {
                double E = c.energy();
                c.last_check_E = E;
                c.last_check_step = c.step;
                khs::write_energy_row(c, c.step, E);
                double ratio = E / c.E0;
                if (std::isnan(E) || std::isinf(E) || ratio > c.cfg.energy_thresh) {
                    if (halted) *halted = 1;
                    break;
                }
            }
        }
    });
}

int khs_get_step(khs_ctx* ctx, int* step, double* time) {
    CTX_OR_FAIL(ctx);
    Ctx& c = *ctx->impl;
    if (step) *step = (int)c.step;
    if (time) *time = c.step * c.cfg.dt;
    return KHS_OK;
}

int khs_energy(khs_ctx* ctx, double* E) {
    CTX_OR_FAIL(ctx);
    if (!E) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    return guard([&] { *E = ctx->impl->energy(); });
}

int khs_last_energy(khs_ctx* ctx, double* E, int* step) {
    CTX_OR_FAIL(ctx);
    if (!E || !step) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    *E = ctx->impl->last_check_E;
    *step = (int)ctx->impl->last_check_step;
    return KHS_OK;
}

int khs_energy_noalloc(const void* const arrays[KHS_N_STATE], int nx, int nz, int precision, void* stream,
                       double* out_dev) {
    if (!arrays || !out_dev) { g_last_error = "khs_energy_noalloc: null argument"; return KHS_ERR_ARG; }
    return guard([&] {
        using namespace khs;
        if (nx < 1 || nz < 1) throw Error(KHS_ERR_ARG, "khs_energy_noalloc: nx and nz must be >= 1");
        if ((long)nx * (long)nz * 8L > 0xFFFFFFF0L)
            throw Error(KHS_ERR_ARG, "khs_energy_noalloc: grid above the library's limit of 2^29 - 2 cells");
        if (precision != KHS_FP64 && precision != KHS_FP32)
            throw Error(KHS_ERR_ARG, "khs_energy_noalloc: precision must be KHS_FP64 or KHS_FP32");
        static const int used[15] = {S_EX1, S_EX2, S_EX3, S_EZ1, S_EZ2, S_EZ3, S_BY1, S_BY2, S_BY3,
                                     S_NA, S_PXA, S_PZA, S_NB, S_PXB, S_PZB};
        for (int id : used)
            if (!arrays[id]) throw Error(KHS_ERR_ARG, "khs_energy_noalloc: array " + std::to_string(id) + " is null");
        const size_t N = (size_t)nx * nz;
        const size_t blocks = std::min<size_t>((N + kRedThreads - 1) / kRedThreads, kRedBlocks);
        if (blocks * 2 > kScratchDoubles)
            throw Error(KHS_ERR_ARG, "khs_energy_noalloc: reduction needs more partial sums than the scratch holds");
        double* partials = scratch_of_current_device();
        KParams kp{};
        kp.nx = nx;
        kp.nz = nz;
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (precision == KHS_FP32) {
            State<float> S{};
            for (int i = 0; i < S_COUNT; ++i) S.a[i] = static_cast<float*>(const_cast<void*>(arrays[i]));
            launch_energy<float>(S, kp, partials, out_dev, st);
        } else {
            State<double> S{};
            for (int i = 0; i < S_COUNT; ++i) S.a[i] = static_cast<double*>(const_cast<void*>(arrays[i]));
            launch_energy<double>(S, kp, partials, out_dev, st);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) throw Error(KHS_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    });
}

int khs_energy0(khs_ctx* ctx, double* E0) {
    CTX_OR_FAIL(ctx);
    if (!E0) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    *E0 = ctx->impl->E0;
    return KHS_OK;
}

int khs_gauss_residual(khs_ctx* ctx, double rms[3], double rel[3]) {
    if (!rms || !rel) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    double g[7];
    const int rc = khs_gauss(ctx, g);
    if (rc != KHS_OK) return rc;
    for (int a = 0; a < 3; ++a) {
        rms[a] = g[a];
        rel[a] = g[a] / g[6];
    }
    return KHS_OK;
}

int khs_gauss(khs_ctx* ctx, double out[7]) {
    CTX_OR_FAIL(ctx);
    if (!out) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (c.emhd) throw Error(KHS_ERR_ARG, "khs_gauss: the colour Gauss law is a YM diagnostic");
        if (c.fp32) khs::launch_gauss<float>(c.state<float>(c.cur), c.kp, c.d_partials, c.d_red, c.stream);
        else khs::launch_gauss<double>(c.state<double>(c.cur), c.kp, c.d_partials, c.d_red, c.stream);
        c.check_launch();
        HIPCHK(hipMemcpyAsync(c.h_red, c.d_red, 7 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        const double Nd = (double)c.N;
        for (int k = 0; k < 6; ++k) out[k] = std::sqrt(c.h_red[k] / Nd);
        out[6] = std::sqrt(c.h_red[6] / Nd) / c.cfg.dx + 1e-30;
    });
}

namespace {
void mode_amplitude(Ctx& c, int kz, double out[2]) {
        if (c.emhd) throw Error(KHS_ERR_ARG, "khs_mode_amplitude: YM model only (Az2, Az3)");
        if (kz < 0) throw Error(KHS_ERR_ARG, "kz must be >= 0");
        if (c.fp32)
            khs::launch_mode_amp<float>((const float*)c.st[c.cur][khs::S_AZ2], (const float*)c.st[c.cur][khs::S_AZ3],
                                        kz, c.kp, c.d_partials, c.d_red, c.stream);
        else
            khs::launch_mode_amp<double>((const double*)c.st[c.cur][khs::S_AZ2],
                                         (const double*)c.st[c.cur][khs::S_AZ3], kz, c.kp, c.d_partials,
                                         c.d_red, c.stream);
        c.check_launch();
        HIPCHK(hipMemcpyAsync(c.h_red, c.d_red, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        const double norm = (double)c.nx * (double)c.nz;
        out[0] = c.h_red[0] / norm;
        out[1] = c.h_red[1] / norm;
}
}  // namespace

int khs_mode_amplitude(khs_ctx* ctx, int kz, double out[2]) {
    CTX_OR_FAIL(ctx);
    if (!out) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    return guard([&] { mode_amplitude(*ctx->impl, kz, out); });
}

int khs_mode_spectrum(khs_ctx* ctx, int id, int kmax, double* out) {
    CTX_OR_FAIL(ctx);
    if (!out) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    return guard([&] {
        Ctx& c = *ctx->impl;
        const int slot = khs::state_slot(c, id);
        if (slot < 0) throw Error(KHS_ERR_ARG, "khs_mode_spectrum: state arrays of the model only");
        if (kmax < 1 || kmax > 32 || 2 * kmax > std::min(c.nx, c.nz))
            throw Error(KHS_ERR_ARG, "kmax must be 1..32 and below half the grid");
        if (c.fp32 && !khs::spectrum_mfma_supported(c.nx, c.nz, kmax))
            throw Error(KHS_ERR_ARG, "FP32 spectrum needs nx % 4 == 0 and nz % 16 == 0");
        const size_t e = c.esz;
        if (!c.spec_mem) {
            const size_t zs = c.fp32 ? (size_t)khs::spectrum_z_slices(c.nx, c.nz) : 1;
            const size_t zt = (size_t)c.nz * 64 * e, xt = (size_t)c.nx * 64 * e, cb = zs * 64 * c.nx * e;
            const size_t pb = (size_t)khs::spectrum_x_blocks(c.nx) * 64 * 64 * 4;   // FP32 x partials
            auto ru = [](size_t v) { return (v + 255) / 256 * 256; };
            HIPCHK(hipMalloc((void**)&c.spec_mem, ru(zt) + ru(xt) + ru(cb) + ru(64 * 64 * e) + ru(pb)));
            c.spec_zt = c.spec_mem;
            c.spec_xt = c.spec_mem + ru(zt);
            c.spec_C = c.spec_mem + ru(zt) + ru(xt);
            c.spec_D = c.spec_mem + ru(zt) + ru(xt) + ru(cb);
            c.spec_P = c.spec_mem + ru(zt) + ru(xt) + ru(cb) + ru(64 * 64 * e);
            c.spec_K = 0;
        }
        if (c.spec_K != kmax) {
            std::vector<float> zf((size_t)c.nz * 64), xf((size_t)c.nx * 64);
            khs::fill_spectrum_table(c.nz, kmax, zf.data());
            khs::fill_spectrum_table(c.nx, kmax, xf.data());
            if (c.fp32) {
                HIPCHK(hipMemcpy(c.spec_zt, zf.data(), zf.size() * 4, hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(c.spec_xt, xf.data(), xf.size() * 4, hipMemcpyHostToDevice));
            } else {
                auto dtab = [&](int n, void* dst) {
                    std::vector<double> t((size_t)n * 64, 0.0);
                    for (int z = 0; z < n; ++z)
                        for (int r = 0; r < 2 * kmax; ++r) {
                            const int m = (r < kmax) ? r + 1 : r - kmax + 1;
                            const double ph = 2.0 * M_PI * (double)((long)m * z % n) / (double)n;
                            t[(size_t)z * 64 + r] = (r < kmax) ? std::cos(ph) : std::sin(ph);
                        }
                    HIPCHK(hipMemcpy(dst, t.data(), t.size() * 8, hipMemcpyHostToDevice));
                };
                dtab(c.nz, c.spec_zt);
                dtab(c.nx, c.spec_xt);
            }
            c.spec_K = kmax;
        }
        std::vector<double> D(64 * 64);
        if (c.fp32) {
            khs::launch_spectrum_mfma((const float*)c.st[c.cur][slot], (const float*)c.spec_zt,
                                      (const float*)c.spec_xt, (float*)c.spec_C, (float*)c.spec_P,
                                      (float*)c.spec_D, c.nx, c.nz, c.stream);
            c.check_launch();
            std::vector<float> Df(64 * 64);
            HIPCHK(hipMemcpyAsync(Df.data(), c.spec_D, Df.size() * 4, hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            for (size_t i = 0; i < D.size(); ++i) D[i] = Df[i];
        } else {
            khs::launch_spectrum_f64((const double*)c.st[c.cur][slot], (const double*)c.spec_zt,
                                     (const double*)c.spec_xt, (double*)c.spec_C, (double*)c.spec_D, c.nx, c.nz,
                                     2 * kmax, c.stream);
            c.check_launch();
            HIPCHK(hipMemcpyAsync(D.data(), c.spec_D, D.size() * 8, hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
        }
        // D[r][s]: r = z row (cos kz | sin kz), s = x column (cos kx | sin kx);
        // F = sum f e^{-i(a + b)} -> Re = CC - SS, Im = -(CS + SC)
        const int K = kmax;
        for (int kz = 0; kz < K; ++kz)
            for (int kx = 0; kx < K; ++kx) {
                const double cc = D[(size_t)kz * 64 + kx], cs = D[(size_t)kz * 64 + K + kx];
                const double sc = D[(size_t)(K + kz) * 64 + kx], ss = D[(size_t)(K + kz) * 64 + K + kx];
                out[2 * ((size_t)kz * K + kx) + 0] = cc - ss;
                out[2 * ((size_t)kz * K + kx) + 1] = -(cs + sc);
            }
    });
}

namespace {
struct CkptHeader {
    char magic[8];
    uint32_t version;
    int32_t model, nx, nz, fp32, nsp, periodic_x, narrays;
    int64_t step;
    double E0, dt, dx, dz;
};
static_assert(sizeof(CkptHeader) == 80, "checkpoint header layout");
constexpr char kCkptMagic[8] = {'K', 'H', 'S', 'C', 'K', 'P', 'T', 0};

uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}

CkptHeader ckpt_header(const khs::Ctx& c) {
    CkptHeader h{};
    std::memcpy(h.magic, kCkptMagic, 8);
    h.version = 1;
    h.model = c.emhd ? KHS_MODEL_EMHD : KHS_MODEL_YM;
    h.nx = c.nx; h.nz = c.nz; h.fp32 = c.fp32 ? 1 : 0; h.nsp = c.nsp;
    h.periodic_x = c.kp.periodic_x; h.narrays = khs::S_COUNT;
    h.step = c.step; h.E0 = c.E0; h.dt = c.kp.dt; h.dx = c.kp.dx; h.dz = c.kp.dz;
    return h;
}
}  // namespace

int khs_checkpoint_write(khs_ctx* ctx, const char* path) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (!path) throw Error(KHS_ERR_ARG, "null path");
        khs::require_idle(c, "khs_checkpoint_write");
        const std::string tmp = std::string(path) + ".tmp";
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) throw Error(KHS_ERR_IO, "cannot open " + tmp);
        const CkptHeader h = ckpt_header(c);
        f.write(reinterpret_cast<const char*>(&h), sizeof(h));
        std::vector<double> a(c.N);
        uint64_t hash = 14695981039346656037ull;
        for (int s = 0; s < khs::S_COUNT; ++s) {
            khs::download(c, c.st[c.cur][s], c.fp32, a.data());
            f.write(reinterpret_cast<const char*>(a.data()), (std::streamsize)(c.N * sizeof(double)));
            hash = fnv1a(hash, a.data(), c.N * sizeof(double));
        }
        f.write(reinterpret_cast<const char*>(&hash), sizeof(hash));
        f.close();
        if (!f) throw Error(KHS_ERR_IO, "write failed: " + tmp);
        if (std::rename(tmp.c_str(), path) != 0) throw Error(KHS_ERR_IO, std::string("cannot rename to ") + path);
    });
}

int khs_checkpoint_read(khs_ctx* ctx, const char* path) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (!path) throw Error(KHS_ERR_ARG, "null path");
        khs::require_idle(c, "khs_checkpoint_read");
        std::ifstream f(path, std::ios::binary);
        if (!f) throw Error(KHS_ERR_IO, std::string("cannot open ") + path);
        CkptHeader h{};
        if (!f.read(reinterpret_cast<char*>(&h), sizeof(h)) || std::memcmp(h.magic, kCkptMagic, 8) != 0 ||
            h.version != 1)
            throw Error(KHS_ERR_IO, std::string(path) + " is not a khs checkpoint");
        const CkptHeader want = ckpt_header(c);
        if (h.model != want.model || h.nx != want.nx || h.nz != want.nz || h.fp32 != want.fp32 ||
            h.nsp != want.nsp || h.periodic_x != want.periodic_x || h.narrays != want.narrays ||
            std::memcmp(&h.dt, &want.dt, sizeof(double)) != 0 || std::memcmp(&h.dx, &want.dx, sizeof(double)) != 0 ||
            std::memcmp(&h.dz, &want.dz, sizeof(double)) != 0)
            throw Error(KHS_ERR_CONFIG, std::string(path) + " was written by a different configuration");
        std::vector<std::vector<double>> a(khs::S_COUNT, std::vector<double>(c.N));
        uint64_t hash = 14695981039346656037ull;
        for (int s = 0; s < khs::S_COUNT; ++s) {
            if (!f.read(reinterpret_cast<char*>(a[s].data()), (std::streamsize)(c.N * sizeof(double))))
                throw Error(KHS_ERR_IO, std::string(path) + " is truncated");
            hash = fnv1a(hash, a[s].data(), c.N * sizeof(double));
        }
        uint64_t stored = 0;
        if (!f.read(reinterpret_cast<char*>(&stored), sizeof(stored)) || stored != hash)
            throw Error(KHS_ERR_IO, std::string(path) + ": checksum mismatch");
        for (int s = 0; s < khs::S_COUNT; ++s) khs::upload_state(c, c.cur, s, a[s].data());
        c.step = h.step;
        c.E0 = h.E0;
    });
}

int khs_write_snapshot(khs_ctx* ctx, const char* dir, int step) {
    CTX_OR_FAIL(ctx);
    if (!dir) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (c.emhd) {
            // emhdgpu2d/main_coupled.cu:286-300: coupled_%05d.csv with X,Z,By,Pz
            std::vector<double> by(c.N), pz(c.N);
            khs::download(c, c.st[c.cur][khs::E_BY], c.fp32, by.data());
            khs::download(c, c.st[c.cur][12 + khs::EV_PZ], c.fp32, pz.data());
            std::ostringstream fn;
            fn << dir << "/coupled_" << std::setfill('0') << std::setw(5) << step << ".csv";
            std::ofstream out(fn.str());
            if (!out) throw Error(KHS_ERR_IO, "cannot write " + fn.str());
            out << "X,Z,By,Pz\n";
            for (int z = 0; z < c.nz; ++z)
                for (int x = 0; x < c.nx; ++x) {
                    size_t i = (size_t)x + (size_t)z * c.nx;
                    out << x * c.cfg.dx << ',' << z * c.cfg.dz << ',' << by[i] << ',' << pz[i] << '\n';
                }
            if (!out) throw Error(KHS_ERR_IO, "write failed: " + fn.str());
            return;
        }
        // column order of main_ym.cu:291-292
        const int cols[24] = {KHS_BY1, KHS_BY2, KHS_BY3, KHS_AZ2, KHS_AZ3, KHS_PZA, KHS_PXA, KHS_Q1A,
                              KHS_PZB, KHS_PXB, KHS_Q1B, KHS_EX1, KHS_EX2, KHS_EX3, KHS_EZ1, KHS_EZ2,
                              KHS_EZ3, KHS_NA, KHS_NB, KHS_Q2A, KHS_Q3A, KHS_Q2B, KHS_Q3B, KHS_AZ1};
        std::vector<std::vector<double>> h(24, std::vector<double>(c.N));
        for (int k = 0; k < 24; ++k) khs::download(c, c.st[c.cur][cols[k]], c.fp32, h[k].data());
        std::ostringstream fn;
        fn << dir << "/ym_" << std::setfill('0') << std::setw(6) << step << ".csv";
        std::ofstream out(fn.str());
        if (!out) throw Error(KHS_ERR_IO, "cannot write " + fn.str());
        out << "X,Z,By1,By2,By3,Az2,Az3,PzA,PxA,Q1A,PzB,PxB,Q1B,"
               "Ex1,Ex2,Ex3,Ez1,Ez2,Ez3,nA,nB,Q2A,Q3A,Q2B,Q3B,Az1\n";
        const double DX = c.cfg.dx, DZ = c.cfg.dz;
        for (int z = 0; z < c.nz; ++z)
            for (int x = 0; x < c.nx; ++x) {
                size_t i = (size_t)x + (size_t)z * c.nx;
                out << x * DX << ',' << z * DZ;
                for (int k = 0; k < 24; ++k) out << ',' << h[k][i];
                out << '\n';
            }
        if (!out) throw Error(KHS_ERR_IO, "write failed: " + fn.str());
    });
}

int khs_energy_csv_open(khs_ctx* ctx, const char* path) {
    CTX_OR_FAIL(ctx);
    if (!path) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    return guard([&] {
        Ctx& c = *ctx->impl;
        if (c.ecsv.is_open()) c.ecsv.close();
        c.ecsv.open(path);
        if (!c.ecsv) throw Error(KHS_ERR_IO, std::string("cannot open ") + path);
        c.ecsv << std::scientific << std::setprecision(8);
        c.ecsv << "step,time,E_total,E_ratio\n";
        c.ecsv << "0,0," << c.E0 << ",1.0\n";
        c.ecsv.flush();
    });
}

int khs_energy_csv_row(khs_ctx* ctx, int step, double E) {
    CTX_OR_FAIL(ctx);
    return guard([&] { khs::write_energy_row(*ctx->impl, step, E); });
}

int khs_energy_csv_close(khs_ctx* ctx) {
    CTX_OR_FAIL(ctx);
    return guard([&] { if (ctx->impl->ecsv.is_open()) ctx->impl->ecsv.close(); });
}

int khs_sync(khs_ctx* ctx) {
    CTX_OR_FAIL(ctx);
    return guard([&] { HIPCHK(hipStreamSynchronize(ctx->impl->stream)); });
}

int khs_timing_enable(khs_ctx* ctx, int on) {
    CTX_OR_FAIL(ctx);
    return guard([&] { ctx->impl->harvest(); ctx->impl->timing = (on != 0); });
}

int khs_timing_reset(khs_ctx* ctx) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        c.harvest();
        for (int i = 0; i < khs::K_NSLOTS; ++i) { c.ms[i] = 0.0; c.cnt[i] = 0; }
    });
}

int khs_kernel_times(khs_ctx* ctx, double* ms, int* launches, int nslots) {
    CTX_OR_FAIL(ctx);
    return guard([&] {
        Ctx& c = *ctx->impl;
        c.harvest();
        for (int i = 0; i < nslots && i < khs::K_NSLOTS; ++i) {
            if (ms) ms[i] = c.ms[i];
            if (launches) launches[i] = c.cnt[i];
        }
    });
}

int khs_get_tuning(khs_ctx* ctx, int out[13]) {
    CTX_OR_FAIL(ctx);
    if (!out) { g_last_error = "null argument"; return KHS_ERR_ARG; }
    Ctx& c = *ctx->impl;
    static const int xs[3] = {64, 128, 256}, xr[3] = {1, 4, 16}, zs[3] = {16, 32, 64};
    out[0] = xs[c.xvar % 3];
    out[1] = xr[(c.xvar / 3) % 3];
    out[2] = zs[c.zvar % 3];
    out[3] = khs::z_prefetch_of(c.zvar);
    out[4] = c.mvar;
    out[5] = c.xvar / 9;
    out[6] = c.xzvar;
    out[7] = khs::z_groups_of(c.zvar);
    out[8] = 0;
    out[9] = c.zmvar;
    out[10] = c.fused ? 1 : 0;
    out[11] = c.mxz ? 1 : 0;
    out[12] = c.mxzvar;
    if (c.xvar >= 18) {
        const int cells = c.xvar == 18 ? 4 : 8;
        out[0] = 62 * cells;
        out[1] = 1;
        out[5] = 1;
        out[8] = cells;
    }
    return KHS_OK;
}

const char* khs_kernel_name(int slot) {
    if (slot < 0 || slot >= khs::K_NSLOTS) return nullptr;
    return khs::kSlotNames[slot];
}

}  // extern "C"

// ==========================================================================
// C ABI: batches
// ==========================================================================
struct khs_batch {
    std::unique_ptr<khs::Batch> impl;
};

extern "C" {

int khs_batch_create(const khs_config* cfgs, int n, int lanes, int segment, khs_batch** out) {
    if (!cfgs || !out || n < 1 || lanes < 1 || segment < 2 || segment % 2) {
        g_last_error = "khs_batch_create: need cfgs, n >= 1, lanes >= 1 and an even segment >= 2";
        return KHS_ERR_ARG;
    }
    *out = nullptr;
    return guard([&] {
        auto b = std::make_unique<khs::Batch>();
        b->seg = segment;
        for (int i = 0; i < n; ++i) b->runs.push_back(khs::make_ctx(cfgs[i]));
        HIPCHK(hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking));
        for (int l = 0; l < std::min(lanes, n); ++l) {
            hipStream_t s;
            HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            b->lanes.push_back(s);
        }
        for (size_t l = 0; l <= b->lanes.size(); ++l) {
            hipEvent_t e;
            HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            b->ev.push_back(e);
        }
        b->active.assign(n, 1);
        b->cap.assign(n, std::numeric_limits<long>::max());
        auto h = new khs_batch;
        h->impl = std::move(b);
        *out = h;
    });
}

int khs_batch_destroy(khs_batch* b) {
    if (!b) return KHS_OK;
    int rc = guard([&] { b->impl.reset(); });
    delete b;
    return rc;
}

int khs_batch_init(khs_batch* b) {
    if (!b || !b->impl) { g_last_error = "null batch"; return KHS_ERR_ARG; }
    return guard([&] {
        khs::Batch& B = *b->impl;
        B.release_graph();
        for (size_t i = 0; i < B.runs.size(); ++i) {
            khs_ctx tmp;
            tmp.impl = std::move(B.runs[i]);
            const int rc = khs_init(&tmp, nullptr);
            B.runs[i] = std::move(tmp.impl);
            if (rc != KHS_OK) throw Error(rc, "run " + std::to_string(i) + ": " + g_last_error);
            B.active[i] = 1;
        }
    });
}

static int batch_run_graphs(khs_batch* b, int nsteps, int* steps_done, int* halted, bool packed) {
    if (!b || !b->impl) { g_last_error = "null batch"; return KHS_ERR_ARG; }
    return guard([&] {
        khs::Batch& B = *b->impl;
        const int n = (int)B.runs.size();
        int left = nsteps;
        while (left > 0) {
            // runs that have neither halted nor reached their step limit
            std::vector<int> set;
            long rem = left;
            for (int i = 0; i < n; ++i)
                if (B.active[i] && B.runs[i]->step < B.cap[i]) {
                    set.push_back(i);
                    rem = std::min(rem, B.cap[i] - B.runs[i]->step);
                }
            if (set.empty()) break;
            int K = (int)std::min<long>(B.seg, rem);
            if (K < 2) {                          // odd remainder: one eager step per run
                for (int i : set) khs::full_step(*B.runs[i]);
                left -= 1;
                continue;
            }
            K -= K % 2;
            if (packed && !B.packable(set))
                throw Error(KHS_ERR_CONFIG, "khs_batch_run_packed: runs must share the YM model and precision and "
                                            "filter only with xi_cut / sponge / suppress_kz0");
            // a graph steps the ping-pong buffers it was captured from: an
            // eager step since (an odd remainder) needs a new capture
            if (!B.exec || B.exec_k != K || B.exec_set != set || B.exec_packed != packed ||
                B.exec_cur != B.cur_of(set)) {
                if (packed) B.capture_packed(K, set);
                else B.capture(K, set);
            }
            HIPCHK(hipGraphLaunch(B.exec, B.origin));
            HIPCHK(hipStreamSynchronize(B.origin));
            for (int i : set) {
                Ctx& c = *B.runs[i];
                c.step += K;
                const double E = c.h_red[0];
                c.last_check_E = E;
                c.last_check_step = c.step;
                khs::write_energy_row(c, c.step, E);
                const double ratio = E / c.E0;
                if (std::isnan(E) || std::isinf(E) || ratio > c.cfg.energy_thresh) B.active[i] = 0;
            }
            left -= K;
        }
        for (int i = 0; i < n; ++i) {
            if (steps_done) steps_done[i] = (int)B.runs[i]->step;
            if (halted) halted[i] = B.active[i] ? 0 : 1;
        }
    });
}

int khs_batch_run(khs_batch* b, int nsteps, int* steps_done, int* halted) {
    return batch_run_graphs(b, nsteps, steps_done, halted, false);
}

int khs_batch_run_packed(khs_batch* b, int nsteps, int* steps_done, int* halted) {
    return batch_run_graphs(b, nsteps, steps_done, halted, true);
}

int khs_batch_run_eager(khs_batch* b, int nsteps, int* steps_done, int* halted) {
    if (!b || !b->impl) { g_last_error = "null batch"; return KHS_ERR_ARG; }
    return guard([&] {
        khs::Batch& B = *b->impl;
        for (size_t i = 0; i < B.runs.size(); ++i) {
            if (!B.active[i]) continue;
            khs_ctx tmp;
            tmp.impl = std::move(B.runs[i]);
            int h = 0;
            const int rc = khs_step(&tmp, nsteps, &h);
            B.runs[i] = std::move(tmp.impl);
            if (rc != KHS_OK) throw Error(rc, "run " + std::to_string(i) + ": " + g_last_error);
            if (h) B.active[i] = 0;
        }
        for (size_t i = 0; i < B.runs.size(); ++i) {
            if (steps_done) steps_done[i] = (int)B.runs[i]->step;
            if (halted) halted[i] = B.active[i] ? 0 : 1;
        }
    });
}

int khs_batch_run_concurrent(khs_batch* b, int nsteps, int threads, int* steps_done, int* halted) {
    if (!b || !b->impl) { g_last_error = "null batch"; return KHS_ERR_ARG; }
    if (threads < 1) { g_last_error = "threads must be >= 1"; return KHS_ERR_ARG; }
    return guard([&] {
        khs::Batch& B = *b->impl;
        const int n = (int)B.runs.size();
        // runs dealt to host threads by longest processing time first (cells x
        // steps to go); each thread steps its runs one after another with
        // khs_step (graph-replayed energy strides) on the run's own stream
        std::vector<std::vector<int>> part(threads);
        std::vector<double> load(threads, 0.0);
        std::vector<int> order;
        std::vector<double> cost(n, 0.0);
        for (int i = 0; i < n; ++i) {
            const Ctx& c = *B.runs[i];
            const long todo = std::min<long>(nsteps, B.cap[i] - c.step);
            if (!B.active[i] || todo <= 0) continue;
            cost[i] = (double)c.nx * (double)c.nz * (double)todo;
            order.push_back(i);
        }
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
        // Everything but replaying happens here, on this thread, before the
        // workers start: the first context of each grid shape times its tile
        // variants (the others take the cached choice), and every run's
        // energy-stride graph is captured and instantiated.  The workers then
        // only launch graphs (and eager steps for a stride's remainder) on
        // their runs' streams; no stream capture runs while another thread
        // launches work.
        for (int i : order) {
            Ctx& c = *B.runs[i];
            c.autotune();
            HIPCHK(hipStreamSynchronize(c.stream));
            const int stride = c.cfg.energy_stride;
            const long todo = std::min<long>(nsteps, B.cap[i] - c.step);
            if (khs::stride_graphs(c) && todo >= stride && c.step % stride == 0) {
                if (!c.sgraph[c.cur]) khs::capture_stride(c, stride);
                if ((stride & 1) && todo >= 2 * stride && !c.sgraph[1 - c.cur]) {
                    // the graph of the next stride, captured from the buffer
                    // and step parity it will start on
                    c.cur ^= 1;
                    c.step += stride;
                    try {
                        khs::capture_stride(c, stride);
                    } catch (...) {
                        c.cur ^= 1;
                        c.step -= stride;
                        throw;
                    }
                    c.cur ^= 1;
                    c.step -= stride;
                }
            }
        }
        for (int i : order) {
            const int t = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            part[t].push_back(i);
            load[t] += cost[i];
        }
        std::vector<int> rc(threads, KHS_OK);
        std::vector<std::string> msg(threads);
        auto work = [&](int t) {
            for (int i : part[t]) {
                khs_ctx tmp;
                tmp.impl = std::move(B.runs[i]);
                const int todo = (int)std::min<long>(nsteps, B.cap[i] - tmp.impl->step);
                int h = 0;
                const int r = khs_step(&tmp, todo, &h);
                B.runs[i] = std::move(tmp.impl);
                if (r != KHS_OK) { rc[t] = r; msg[t] = "run " + std::to_string(i) + ": " + g_last_error; return; }
                if (h) B.active[i] = 0;
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < threads; ++t) pool.emplace_back(work, t);
        work(0);
        for (auto& th : pool) th.join();
        for (int t = 0; t < threads; ++t)
            if (rc[t] != KHS_OK) throw Error(rc[t], msg[t]);
        for (int i = 0; i < n; ++i) {
            if (steps_done) steps_done[i] = (int)B.runs[i]->step;
            if (halted) halted[i] = B.active[i] ? 0 : 1;
        }
    });
}

int khs_batch_get_array(khs_batch* b, int run, int id, double* host, size_t n) {
    if (!b || !b->impl || run < 0 || run >= (int)b->impl->runs.size()) {
        g_last_error = "bad batch or run index";
        return KHS_ERR_ARG;
    }
    khs_ctx tmp;
    tmp.impl = std::move(b->impl->runs[run]);
    const int rc = khs_get_array(&tmp, id, host, n);
    b->impl->runs[run] = std::move(tmp.impl);
    return rc;
}

int khs_batch_set_array(khs_batch* b, int run, int id, const double* host, size_t n) {
    if (!b || !b->impl || run < 0 || run >= (int)b->impl->runs.size()) {
        g_last_error = "bad batch or run index";
        return KHS_ERR_ARG;
    }
    khs_ctx tmp;
    tmp.impl = std::move(b->impl->runs[run]);
    const int rc = khs_set_array(&tmp, id, host, n);
    b->impl->runs[run] = std::move(tmp.impl);
    return rc;
}

int khs_batch_set_step_limit(khs_batch* b, int run, int max_step) {
    if (!b || !b->impl || run < 0 || run >= (int)b->impl->runs.size()) {
        g_last_error = "bad batch or run index";
        return KHS_ERR_ARG;
    }
    b->impl->cap[run] = max_step > 0 ? (long)max_step : std::numeric_limits<long>::max();
    return KHS_OK;
}

int khs_batch_mode_amplitude(khs_batch* b, int run, int kz, double out[2]) {
    if (!b || !b->impl || run < 0 || run >= (int)b->impl->runs.size() || !out) {
        g_last_error = "bad batch, run index or null argument";
        return KHS_ERR_ARG;
    }
    return guard([&] { mode_amplitude(*b->impl->runs[run], kz, out); });
}

int khs_batch_energy(khs_batch* b, int run, double* E, double* E0) {
    if (!b || !b->impl || run < 0 || run >= (int)b->impl->runs.size()) {
        g_last_error = "bad batch or run index";
        return KHS_ERR_ARG;
    }
    return guard([&] {
        Ctx& c = *b->impl->runs[run];
        if (E) *E = c.energy();
        if (E0) *E0 = c.E0;
    });
}

}  // extern "C"
