// Host-side LP of the SLP step (Implementation/StrSLP.m:221-279):
//
//     min c's   s.t.  a's <= r  (or a's = r),   sL <= s <= sU,
//
// one linear row plus box bounds. This is a continuous knapsack: for a
// multiplier lam the Lagrangian minimiser over the box is s_i = sL_i where
// c_i + lam*a_i > 0 and s_i = sU_i where it is < 0, and a's(lam) is
// non-increasing in lam. The optimum sits at the breakpoint lam* = -c_k/a_k
// where a's(lam) crosses r; only elements whose breakpoint equals lam* lie
// between their bounds. lam* is the multiplier linprog reports as
// lambda.ineqlin (lambda.eqlin for the equality form), which the reference
// uses in the Lagrangian gradient (Implementation/StrSLP.m:480-485).
//
// lam* is found by a weighted selection over the breakpoints: histogram
// passes over an order-preserving integer image of the breakpoints narrow the
// candidate range 65536-fold each, so the work is a few streaming passes over
// the inputs, split across host threads.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "tmg.h"

namespace {

constexpr int NBIN = 65536;

// Order-preserving map of doubles to unsigned integers (total order for
// non-NaN values, -0 below +0).
inline uint64_t okey(double v) {
    uint64_t u;
    std::memcpy(&u, &v, sizeof u);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

inline double lo_state(double a, double sl, double su) { return a > 0 ? su : sl; }
inline double hi_state(double a, double sl, double su) { return a > 0 ? sl : su; }

int lp_threads(int64_t n) {
    int t = (int)std::thread::hardware_concurrency();
    if (const char *e = std::getenv("TMG_LP_THREADS")) t = std::atoi(e);
    t = std::max(1, std::min(t, 16));
    return (int)std::max<int64_t>(1, std::min<int64_t>(t, n / 65536 + 1));
}

template <class F>
void par_chunks(int nt, int64_t n, F &&f) {
    if (nt == 1) { f(0, (int64_t)0, n); return; }
    std::vector<std::thread> th;
    th.reserve((size_t)nt);
    for (int k = 0; k < nt; ++k) {
        const int64_t b = n * k / nt, e = n * (k + 1) / nt;
        th.emplace_back([&f, k, b, e] { f(k, b, e); });
    }
    for (auto &t : th) t.join();
}

struct Cand {
    uint64_t k;
    double w;
};

}  // namespace

extern "C" int tmg_lp_volume(int64_t n, const double *c, const double *a, const double *sL, const double *sU,
                             double r, int equality, double *s, double *lambda) {
    if (n < 0 || (n > 0 && (!c || !a || !sL || !sU || !s)) || !lambda || !std::isfinite(r)) return TMG_ERR_ARG;
    const int nt = lp_threads(n);
    // For the inequality form only lam >= 0 is admissible: breakpoints at or
    // below 0 are already crossed at lam = 0+.
    const double lam_min = equality ? -INFINITY : 0.0;

    // Pass 1: validate; a's with every element in its high-lam state; total
    // weight and key range of the admissible breakpoints.
    struct P1 {
        int bad = 0;
        double hinf = 0, above = 0;
        uint64_t kmin = ~0ull, kmax = 0;
    };
    std::vector<P1> p1((size_t)nt);
    par_chunks(nt, n, [&](int k, int64_t b, int64_t e) {
        P1 q;
        for (int64_t i = b; i < e; ++i) {
            if (!(sL[i] <= sU[i]) || !std::isfinite(c[i]) || !std::isfinite(a[i]) || !std::isfinite(sL[i]) ||
                !std::isfinite(sU[i])) { q.bad = 1; break; }
            if (a[i] == 0.0) continue;
            q.hinf += a[i] * hi_state(a[i], sL[i], sU[i]);
            const double w = std::fabs(a[i]) * (sU[i] - sL[i]);
            const double t = -c[i] / a[i];
            if (w > 0.0 && t > lam_min) {
                q.above += w;
                const uint64_t kk = okey(t);
                q.kmin = std::min(q.kmin, kk);
                q.kmax = std::max(q.kmax, kk);
            }
        }
        p1[(size_t)k] = q;
    });
    double hinf = 0, above = 0;
    uint64_t kmin = ~0ull, kmax = 0;
    for (const P1 &q : p1) {
        if (q.bad) return TMG_ERR_ARG;
        hinf += q.hinf; above += q.above;
        kmin = std::min(kmin, q.kmin); kmax = std::max(kmax, q.kmax);
    }
    const double R = r - hinf;   // weight that may stay in the low-lam state

    int rc = TMG_OK;
    double lam;
    bool have_key = false;
    uint64_t kstar = 0;
    if (R < 0.0) {
        rc = TMG_NOT_CONVERGED;   // even the smallest a's exceeds r: return that point
        lam = INFINITY;
    } else if (above <= R) {
        // the row is slack at lam_min (inequality: lam = 0)
        lam = 0.0;
        if (equality) {
            if (above < R) { rc = TMG_NOT_CONVERGED; lam = -INFINITY; }   // r above the largest a's
            else if (kmin <= kmax) { have_key = true; kstar = kmin; }      // lam = smallest breakpoint
        }
    } else {
        // Select the largest key K with W(>K) <= R < W(>=K), W = weight of
        // breakpoints above. First level streams the inputs; deeper levels work
        // on the gathered candidates of the chosen bin.
        double acc = 0.0;   // weight of breakpoints above the current key range
        bool slack = false;
        uint64_t lo = kmin, hi = kmax;
        std::vector<Cand> cand;
        bool first = true;
        for (;;) {
            const uint64_t span = hi - lo;
            const int shift = span >= (uint64_t)NBIN ? 64 - __builtin_clzll(span / NBIN + 1) : 0;
            auto bin = [&](uint64_t kk) { return (size_t)((kk - lo) >> shift); };
            const size_t nb = (size_t)(span >> shift) + 1;
            std::vector<double> hist(nb, 0.0);
            if (first) {
                std::vector<std::vector<double>> th((size_t)nt, std::vector<double>(nb, 0.0));
                par_chunks(nt, n, [&](int k, int64_t b, int64_t e) {
                    double *h = th[(size_t)k].data();
                    for (int64_t i = b; i < e; ++i) {
                        if (a[i] == 0.0) continue;
                        const double w = std::fabs(a[i]) * (sU[i] - sL[i]);
                        const double t = -c[i] / a[i];
                        if (w > 0.0 && t > lam_min) h[bin(okey(t))] += w;
                    }
                });
                for (auto &h : th)
                    for (size_t j = 0; j < nb; ++j) hist[j] += h[j];
            } else {
                for (const Cand &q : cand) hist[bin(q.k)] += q.w;
            }
            // scan bins from the top until the cumulative weight passes R
            size_t jb = 0;
            bool found = false;
            for (size_t j = nb; j-- > 0;) {
                if (acc + hist[j] > R) { jb = j; found = true; break; }
                acc += hist[j];
            }
            if (!found) {
                // rounding put the total weight at or below R: the row is slack
                kstar = kmin;
                slack = true;
                break;
            }
            const uint64_t blo = lo + ((uint64_t)jb << shift);
            const uint64_t bhi = shift ? std::min<uint64_t>(hi, blo + ((1ull << shift) - 1)) : blo;
            if (shift == 0) { kstar = blo; break; }
            // gather the chosen bin's candidates
            std::vector<Cand> next;
            if (first) {
                std::vector<std::vector<Cand>> th((size_t)nt);
                par_chunks(nt, n, [&](int k, int64_t b, int64_t e) {
                    auto &v = th[(size_t)k];
                    for (int64_t i = b; i < e; ++i) {
                        if (a[i] == 0.0) continue;
                        const double w = std::fabs(a[i]) * (sU[i] - sL[i]);
                        const double t = -c[i] / a[i];
                        if (!(w > 0.0 && t > lam_min)) continue;
                        const uint64_t kk = okey(t);
                        if (kk >= blo && kk <= bhi) v.push_back({kk, w});
                    }
                });
                for (auto &v : th) next.insert(next.end(), v.begin(), v.end());
            } else {
                for (const Cand &q : cand)
                    if (q.k >= blo && q.k <= bhi) next.push_back(q);
            }
            cand.swap(next);
            first = false;
            lo = blo; hi = bhi;
            uint64_t cmin = ~0ull, cmax = 0;
            for (const Cand &q : cand) { cmin = std::min(cmin, q.k); cmax = std::max(cmax, q.k); }
            lo = cmin; hi = cmax;
        }
        have_key = !(slack && !equality);
        if (!have_key) lam = 0.0;
    }
    if (have_key) {
        // back from the key to the breakpoint value
        uint64_t u = (kstar >> 63) ? (kstar & 0x7fffffffffffffffull) : ~kstar;
        std::memcpy(&lam, &u, sizeof lam);
    }

    // Fill s: breakpoints strictly above lam take the low state, strictly
    // below the high state; ties at lam share what is left of R in index order.
    std::vector<double> tie((size_t)nt, 0.0), above_w((size_t)nt, 0.0);
    const bool finite = std::isfinite(lam);
    if (finite)
        par_chunks(nt, n, [&](int k, int64_t b, int64_t e) {
            double tw = 0, aw = 0;
            for (int64_t i = b; i < e; ++i) {
                if (a[i] == 0.0) continue;
                const double t = -c[i] / a[i];
                const double w = std::fabs(a[i]) * (sU[i] - sL[i]);
                if (t > lam) aw += w;
                else if (t == lam) tw += w;
            }
            tie[(size_t)k] = tw;
            above_w[(size_t)k] = aw;
        });
    std::vector<double> rem0((size_t)nt, 0.0);
    double rem = R;
    for (int k = 0; k < nt; ++k) rem -= above_w[(size_t)k];
    for (int k = 0; k < nt; ++k) {
        rem0[(size_t)k] = rem;
        rem = std::max(0.0, rem - tie[(size_t)k]);
    }
    par_chunks(nt, n, [&](int k, int64_t b, int64_t e) {
        double rm = rem0[(size_t)k];
        for (int64_t i = b; i < e; ++i) {
            if (a[i] == 0.0) {
                s[i] = c[i] > 0 ? sL[i] : (c[i] < 0 ? sU[i] : std::min(std::max(0.0, sL[i]), sU[i]));
                continue;
            }
            const double t = -c[i] / a[i];
            const double lo_s = lo_state(a[i], sL[i], sU[i]), hi_s = hi_state(a[i], sL[i], sU[i]);
            if (t > lam) {
                s[i] = lo_s;
            } else if (t < lam || !finite) {
                s[i] = hi_s;
            } else {
                const double w = std::fabs(a[i]) * (sU[i] - sL[i]);
                if (rm >= w) { s[i] = lo_s; rm -= w; }
                else if (rm > 0.0 && w > 0.0) { s[i] = hi_s + (rm / w) * (lo_s - hi_s); rm = 0.0; }
                else s[i] = hi_s;
            }
        }
    });
    *lambda = finite ? lam : 0.0;
    return rc;
}
