// sp_host.h -- internal host-side declarations of the engine.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/stophys_gpu.h"
#include "sp_types.h"

namespace sp {

int request_dims(int problem, uint32_t* dy, uint32_t* dw, uint32_t* dout);
// the parameter rows of every particle (param-major, row stride `stride`), split over threads for large ensembles
void compute_params(const sp_request& rq, const sp_particles& pt, double* out, uint64_t stride);
// the same for particles [p0, p1)
void compute_params_range(const sp_request& rq, const sp_particles& pt, double* out, uint64_t stride, uint64_t p0,
                          uint64_t p1);
void time_and_field(const sp_request& rq, double* times, double* fields);
// the request's field as FieldParams (angular frequencies, rectangular edge constants)
FieldParams field_params(const sp_request& rq);

// fn(begin, end) over [0, n) in contiguous ranges, on up to 16 threads (one range per 65,536 items at least); every
// item is handled by exactly one call, so results written per item do not depend on the split.  A thread that
// cannot be started runs its range on the calling thread.
template <class F> void parallel_ranges(uint64_t n, F&& fn) {
    const uint64_t hw = std::max(1u, std::thread::hardware_concurrency());
    const uint64_t nt = std::min<uint64_t>(std::min<uint64_t>(16, hw), (n >> 16) + 1);
    if (nt <= 1) { fn(uint64_t(0), n); return; }
    const uint64_t chunk = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (uint64_t t = 1; t < nt; ++t) {
        const uint64_t a = std::min(n, t * chunk), b = std::min(n, a + chunk);
        if (a >= b) continue;
        try {
            th.emplace_back([&fn, a, b] { fn(a, b); });
        } catch (...) {
            fn(a, b);
        }
    }
    fn(uint64_t(0), std::min(n, chunk));
    for (auto& x : th) x.join();
}

// thread-local error message for sp_last_error()
int set_error(int code, const std::string& msg);

}  // namespace sp
