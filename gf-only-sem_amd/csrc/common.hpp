// common.hpp -- shared host helpers: error state, GLL setup, parallel loops.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "gfx.h"

namespace gfx {

// Thread-local last-error message behind gfx_last_error().
void set_error(const std::string& msg);
void clear_error();

struct Error : std::runtime_error {
    gfx_status code;
    Error(gfx_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Runs fn and converts exceptions into a status + message.
template <typename Fn>
gfx_status guarded(Fn&& fn) {
    try {
        clear_error();
        fn();
        return GFX_OK;
    } catch (const Error& e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return GFX_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return GFX_ERR_INVALID;
    }
}

inline void require(bool ok, const std::string& msg) {
    if (!ok) throw Error(GFX_ERR_INVALID, msg);
}

// GLL points, weights and derivative matrix D[i][j] = l_j'(x_i) for order N.
// Same construction as forward/share/include/gf/gll.hpp:48-131 (Newton on
// P_N' from Chebyshev-Gauss-Lobatto guesses; closed-form weights and D).
struct GLL {
    int N = 0, n = 0;
    double x[16] = {0}, w[16] = {0}, D[256] = {0};
    explicit GLL(int order);
    void lagrange(double xi, double* ell) const;
};

// Static-partitioned parallel loop over [0, n) with up to `threads` workers.
template <typename Fn>
void parallel_for(int64_t n, Fn&& fn, int threads = 0) {
    if (threads <= 0) {
        unsigned hc = std::thread::hardware_concurrency();
        threads = (int)std::min<unsigned>(hc ? hc : 1, 16);
    }
    if (n < 4096 || threads == 1) {
        for (int64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([b, e, &fn] {
            for (int64_t i = b; i < e; ++i) fn(i);
        });
    }
    for (auto& th : pool) th.join();
}

}  // namespace gfx
