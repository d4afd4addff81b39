// khs_analysis.cpp — growth-rate fitters of the analysis pipeline, host side.
//
//   khs_fit_growth_max_r2   ymgpu2d/analysis/batch_analyze.py:44-69
//                            (same algorithm as dispersion_ym.py): the window
//                            [i, j) of log-amplitude samples, at least
//                            max(4, floor(n_w * min_win_frac)) long, whose
//                            least-squares line has the largest R^2 among
//                            windows with slope >= 1e-6.
//   khs_fit_growth_plateau  batch_analyze.py:72-152: local slopes over a
//                            sliding window of width win_tu, the longest run
//                            of positive slopes whose spread stays within
//                            flat_tol of its median, fitted once more.
//
// Both take the (time, amplitude) series that khs_mode_amplitude produces
// every export (remote_timeseries.py:62-77 computes the same amplitude from
// CSV snapshots).  Straight-line fits use the closed-form least-squares
// solution on centred data.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "../../include/khs.h"

namespace {

struct Line {
    double slope, icpt;
};

// Least-squares line through (x[k], y[k]), k in [a, b).
Line fit_line(const double* x, const double* y, int a, int b) {
    const int n = b - a;
    double mx = 0.0, my = 0.0;
    for (int k = a; k < b; ++k) { mx += x[k]; my += y[k]; }
    mx /= n; my /= n;
    double sxx = 0.0, sxy = 0.0;
    for (int k = a; k < b; ++k) {
        const double dx = x[k] - mx;
        sxx += dx * dx;
        sxy += dx * (y[k] - my);
    }
    const double s = (sxx > 0.0) ? sxy / sxx : 0.0;
    return {s, my - s * mx};
}

double median(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    return (n % 2) ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

}  // namespace

extern "C" {

int khs_fit_growth_max_r2(const double* t, const double* amp, int n, double t_min, double t_max,
                          double min_win_frac, double out[4]) {
    if (!t || !amp || !out || n < 0) return KHS_ERR_ARG;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    std::vector<double> tw, aw;
    for (int k = 0; k < n; ++k) {
        if (!std::isnan(t_min) && t[k] < t_min) continue;
        if (!std::isnan(t_max) && t[k] > t_max) continue;
        tw.push_back(t[k]);
        aw.push_back(std::log(std::max(amp[k], 1e-30)));
    }
    out[0] = 0.0; out[1] = nan; out[2] = nan; out[3] = 0.0;
    const int m = (int)tw.size();
    if (m < 4) return KHS_OK;
    const int min_win = std::max(4, (int)(m * min_win_frac));
    double best_r2 = -1e9;
    for (int i = 0; i < m; ++i)
        for (int j = i + min_win; j <= m; ++j) {
            const Line c = fit_line(tw.data(), aw.data(), i, j);
            if (c.slope < 1e-6) continue;
            double ss_res = 0.0, mean = 0.0, ss_tot = 0.0;
            for (int k = i; k < j; ++k) mean += aw[k];
            mean /= (j - i);
            for (int k = i; k < j; ++k) {
                const double r = aw[k] - (c.slope * tw[k] + c.icpt);
                ss_res += r * r;
                ss_tot += (aw[k] - mean) * (aw[k] - mean);
            }
            const double r2 = 1.0 - ss_res / (ss_tot + 1e-30);
            if (r2 > best_r2) {
                best_r2 = r2;
                out[0] = c.slope; out[1] = tw[i]; out[2] = tw[j - 1]; out[3] = r2;
            }
        }
    if (best_r2 == -1e9) out[3] = -1e9;
    return KHS_OK;
}

int khs_fit_growth_plateau(const double* t, const double* amp, int n, double win_tu, double min_plateau_tu,
                           double flat_tol, double out[5]) {
    if (!t || !amp || !out || n < 0) return KHS_ERR_ARG;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    out[0] = nan; out[1] = nan; out[2] = nan; out[3] = nan; out[4] = 0.0;
    if (n < 8) return KHS_OK;
    std::vector<double> la(n);
    for (int k = 0; k < n; ++k) la[k] = std::log(std::max(amp[k], 1e-30));
    // local slope series (_local_slope_series, batch_analyze.py:72-87)
    const double half = win_tu / 2.0;
    std::vector<double> tc, sl;
    for (int i = 0; i < n; ++i) {
        std::vector<double> xs, ys;
        for (int k = 0; k < n; ++k)
            if (t[k] >= t[i] - half && t[k] <= t[i] + half) { xs.push_back(t[k]); ys.push_back(la[k]); }
        if (xs.size() < 4) continue;
        tc.push_back(t[i]);
        sl.push_back(fit_line(xs.data(), ys.data(), 0, (int)xs.size()).slope);
    }
    const int nc = (int)tc.size();
    if (nc < 4) return KHS_OK;
    int best_a = -1, best_b = -1;
    double best_dur = 0.0;
    int i = 0;
    while (i < nc) {
        if (!(sl[i] > 1e-6)) { ++i; continue; }
        int a = i, b = i;                 // run = [a, b]
        int j = i + 1;
        while (j < nc && sl[j] > 1e-6) {
            std::vector<double> vals(sl.begin() + a, sl.begin() + j + 1);
            const double med = median(vals);
            const double spread = (*std::max_element(vals.begin(), vals.end()) -
                                   *std::min_element(vals.begin(), vals.end())) / med;
            if (spread <= flat_tol) { b = j; ++j; }
            else break;
        }
        const double dur = tc[b] - tc[a];
        if (dur >= min_plateau_tu && (best_a < 0 || dur > best_dur)) { best_a = a; best_b = b; best_dur = dur; }
        i = (b > a) ? b + 1 : i + 1;
    }
    if (best_a < 0) return KHS_OK;
    const double t0 = tc[best_a], t1 = tc[best_b];
    std::vector<double> xs, ys;
    for (int k = 0; k < n; ++k)
        if (t[k] >= t0 && t[k] <= t1) { xs.push_back(t[k]); ys.push_back(la[k]); }
    out[0] = fit_line(xs.data(), ys.data(), 0, (int)xs.size()).slope;
    out[1] = t0; out[2] = t1; out[3] = best_dur; out[4] = 1.0;
    return KHS_OK;
}

}  // extern "C"
