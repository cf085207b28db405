// Host-side LP of the SLP step (Implementation/StrSLP.m:221-279):
//
//     min c's   s.t.  a's <= r  (or a's = r),   sL <= s <= sU,
//
// one linear row plus box bounds. This is a continuous knapsack: for a
// multiplier lam the Lagrangian minimiser over the box is s_i = sL_i where
// c_i + lam*a_i > 0 and s_i = sU_i where it is < 0, and a's(lam) is
// non-increasing in lam. The optimum is found by locating the breakpoint
// lam* = -c_k/a_k at which a's(lam) crosses r (a weighted selection, expected
// O(n) with nth_element), leaving at most the elements tied at lam* between
// their bounds. lam* is the multiplier linprog reports as lambda.ineqlin
// (lambda.eqlin for the equality form), which the reference uses in the
// Lagrangian gradient (Implementation/StrSLP.m:480-485).
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "tmg.h"

namespace {

struct Brk {
    double t;   // breakpoint -c_i/a_i
    double w;   // drop of a's when lam crosses t: |a_i| (sU_i - sL_i)
};

// State of s_i for lam below / above its breakpoint.
inline double lo_state(double a, double sl, double su) { return a > 0 ? su : sl; }
inline double hi_state(double a, double sl, double su) { return a > 0 ? sl : su; }

}  // namespace

extern "C" int tmg_lp_volume(int64_t n, const double *c, const double *a, const double *sL, const double *sU,
                             double r, int equality, double *s, double *lambda) {
    if (n < 0 || (n > 0 && (!c || !a || !sL || !sU || !s)) || !lambda || !std::isfinite(r)) return TMG_ERR_ARG;
    for (int64_t i = 0; i < n; ++i)
        if (!(sL[i] <= sU[i]) || !std::isfinite(c[i]) || !std::isfinite(a[i])) return TMG_ERR_ARG;

    // a's with every element in its high-lam state, and the breakpoints.
    double hinf = 0.0;
    std::vector<Brk> br;
    br.reserve((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        if (a[i] == 0.0) continue;
        hinf += a[i] * hi_state(a[i], sL[i], sU[i]);
        const double w = std::fabs(a[i]) * (sU[i] - sL[i]);
        if (w > 0.0) br.push_back({-c[i] / a[i], w});
    }
    // For the inequality form only lam >= 0 is admissible: breakpoints at or
    // below 0 are already crossed at lam = 0+.
    const double lam_min = equality ? -INFINITY : 0.0;
    double R = r - hinf;   // weight that may stay in the low-lam state
    double above = 0.0;    // weight of breakpoints > lam_min
    size_t m = 0;
    for (size_t j = 0; j < br.size(); ++j)
        if (br[j].t > lam_min) { br[m++] = br[j]; above += br[j].w; }
    br.resize(m);

    int rc = TMG_OK;
    double lam;
    if (R < 0.0) {
        // even the smallest a's exceeds r: infeasible; return that point
        rc = TMG_NOT_CONVERGED;
        lam = INFINITY;
    } else if (above <= R) {
        // the constraint is slack at lam_min (inequality: lam = 0)
        if (equality && above < R) rc = TMG_NOT_CONVERGED;   // r above max a's
        lam = equality ? (m ? std::min_element(br.begin(), br.end(), [](const Brk &x, const Brk &y) {
                                   return x.t < y.t; })->t
                            : 0.0)
                       : 0.0;
        if (equality && above < R) lam = -INFINITY;
    } else {
        // first breakpoint, in decreasing t, at which the cumulative weight passes R
        size_t lo = 0, hi = m;
        double acc = 0.0;
        auto desc = [](const Brk &x, const Brk &y) { return x.t > y.t; };
        while (hi - lo > 64) {
            const size_t mid = lo + (hi - lo) / 2;
            std::nth_element(br.begin() + (std::ptrdiff_t)lo, br.begin() + (std::ptrdiff_t)mid, br.begin() + (std::ptrdiff_t)hi, desc);
            double wl = 0.0;
            for (size_t j = lo; j < mid; ++j) wl += br[j].w;
            if (acc + wl > R) hi = mid;
            else { acc += wl; lo = mid; }
        }
        std::sort(br.begin() + (std::ptrdiff_t)lo, br.begin() + (std::ptrdiff_t)hi, desc);
        lam = br[hi - 1].t;
        for (size_t j = lo; j < hi; ++j) {
            acc += br[j].w;
            if (acc > R) { lam = br[j].t; break; }
        }
    }

    // Fill s: strictly above lam -> low state, strictly below -> high state,
    // ties at lam share the remaining weight in index order.
    double rem = R;
    if (std::isfinite(lam))
        for (int64_t i = 0; i < n; ++i) {
            if (a[i] == 0.0) continue;
            const double t = -c[i] / a[i];
            if (t > lam) rem -= std::fabs(a[i]) * (sU[i] - sL[i]);
        }
    for (int64_t i = 0; i < n; ++i) {
        if (a[i] == 0.0) {
            s[i] = c[i] > 0 ? sL[i] : (c[i] < 0 ? sU[i] : std::min(std::max(0.0, sL[i]), sU[i]));
            continue;
        }
        const double t = -c[i] / a[i];
        const double lo_s = lo_state(a[i], sL[i], sU[i]), hi_s = hi_state(a[i], sL[i], sU[i]);
        if (t > lam) {
            s[i] = lo_s;
        } else if (t < lam || !std::isfinite(lam)) {
            s[i] = hi_s;
        } else {
            const double w = std::fabs(a[i]) * (sU[i] - sL[i]);
            if (rem >= w) { s[i] = lo_s; rem -= w; }
            else if (rem > 0.0 && w > 0.0) { s[i] = hi_s + (rem / w) * (lo_s - hi_s); rem = 0.0; }
            else s[i] = hi_s;
        }
    }
    *lambda = std::isfinite(lam) ? lam : 0.0;
    return rc;
}
