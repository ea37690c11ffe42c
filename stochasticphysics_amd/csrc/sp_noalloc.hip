// sp_noalloc.hip -- sp_execute_noalloc: one request on caller-owned device buffers, allocating nothing.
//
// For callers that cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
// forbid it).  The caller passes the inputs (parameter block, start state) and the output (the ensemble means)
// as device pointers; everything else the kernels need -- the request descriptor, the per-tile partial sums of
// the output windows, the weight sums and the Newton failure counter -- lives in `scratch`, a file-scope device
// array of SCRATCH_BYTES that each GPU holds from the first call on it until the process ends.  Every call on a
// GPU uses that GPU's one array: calls that overlap on one GPU (on different streams or threads) race.
//
// The work is the plain path's (the same simulation kernels, the same fixed-order reductions, the same weight
// sums), so the means equal those of sp_run / sp_execute bit for bit; only the chunk length is bounded by the
// array (chunking does not change the result).
#include "sp_engine.h"
#include "sp_kernels_impl.h"

namespace sp {

constexpr size_t SCRATCH_BYTES = 32ull << 20;
constexpr size_t SCRATCH_HEADER = 4096;        // descriptor, counter and weight sums; the partial sums follow
constexpr size_t OFF_COUNTER = 512, OFF_WSUM = 768;
static_assert(sizeof(DevRequest) <= OFF_COUNTER, "descriptor does not fit the scratch header");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// device address of `scratch` per device, looked up once per device with that device current
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[64] = {};

static int scratch_on(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= 64) return set_error(SP_ERR_INVALID_ARG, "device index out of range");
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(SP_ERR_DEVICE, std::string("scratch array: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[device] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[device];
    return SP_OK;
}

// the request descriptor into the scratch header, the Newton counter cleared, and the weight sums of
// MeanSimulationResult::normalize (MeanSimulationResult.h:63-75) in the fixed order of weight_sum_fixed
// (one block of WSUM_THREADS threads)
__global__ __launch_bounds__(WSUM_THREADS) void k_noalloc_setup(DevRequest dr, unsigned char* base, const double* mu,
                                                                uint64_t n, uint32_t dout) {
    __shared__ double part[WSUM_THREADS];
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        *reinterpret_cast<DevRequest*>(base) = dr;
        *reinterpret_cast<unsigned long long*>(base + OFF_COUNTER) = 0ull;
    }
    double s = 0.0;
    for (uint64_t p = t; p < n; p += WSUM_THREADS) s += mu[p];
    part[t] = s;
    __syncthreads();
    for (uint32_t h = WSUM_THREADS / 2; h > 0; h >>= 1) {
        if (t < h) part[t] += part[t + h];
        __syncthreads();
    }
    if (t < dout) {
        const bool unit = dout == 6 && t < 3;         // BrownAndNeelRelaxation.h:152-159: the axis has weight 1
        reinterpret_cast<double*>(base + OFF_WSUM)[t] = unit ? (double)n : part[0];
    }
}

__global__ __launch_bounds__(256) void k_noalloc_reduce(ReduceWork rw) { reduce_rows(rw); }

// sums -> means (MeanSimulationResult::normalize); all-zero weights leave the sums, as the host path does
__global__ void k_noalloc_normalize(double* mean, uint64_t m, uint32_t dout, const double* wsum) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    bool zero = true;
    for (uint32_t c = 0; c < dout; ++c) zero = zero && wsum[c] == 0.0;
    if (!zero) mean[i] = mean[i] / wsum[i % dout];
}

}  // namespace sp

using namespace sp;

extern "C" int sp_execute_noalloc(int device, void* stream, const sp_request* req, const double* d_params,
                                  double* d_state, double* d_mean) {
    if (!req || !d_params || !d_state || !d_mean) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    if (int e = validate(req, nullptr, true)) return e;
    if (req->noise != SP_NOISE_PHILOX)
        return set_error(SP_ERR_UNSUPPORTED, "sp_execute_noalloc: replayed noise needs a device buffer (use sp_run)");
    if (req->save_interval)
        return set_error(SP_ERR_UNSUPPORTED, "sp_execute_noalloc: single-simulation outputs need a device buffer (use sp_run)");
    HostRequest hr;
    prepare_shape(req, hr);
    // chunks of at most 4,000 steps (a multiple of lcm(Oversampling, 4)) whose partial sums fit the array
    const uint64_t OS = hr.rq.oversampling, align = lcm64(OS, 4);
    const uint64_t row_bytes = (uint64_t)hr.dout * hr.n_tiles * 8;
    const uint64_t cap_rows = (SCRATCH_BYTES - SCRATCH_HEADER) / row_bytes;
    if (cap_rows < align / OS + 1)
        return set_error(SP_ERR_INVALID_ARG,
                         "sp_execute_noalloc: " + std::to_string(hr.N) + " particles with Oversampling " + std::to_string(OS) +
                             " need " + std::to_string((align / OS + 1) * row_bytes) + " bytes of partial sums per chunk, "
                             "more than the " + std::to_string(SCRATCH_BYTES - SCRATCH_HEADER) + "-byte scratch array holds");
    uint64_t C = std::max<uint64_t>(align, 4000 / align * align);
    C = std::min<uint64_t>(C, (cap_rows - 1) * OS / align * align);
    C = std::min<uint64_t>(C, std::max<uint64_t>(hr.total_steps, align));
    const hipStream_t s = (hipStream_t)stream;

    int prev = 0;
    SP_HIP(hipGetDevice(&prev));
    SP_HIP(hipSetDevice(device));
    struct Restore { int d; ~Restore() { hipSetDevice(d); } } restore{prev};
    unsigned char* base = nullptr;
    if (int e = scratch_on(device, &base)) return e;
    double* partial = (double*)(base + SCRATCH_HEADER);
    const double* wsum = (const double*)(base + OFF_WSUM);

    ReqDev d;
    d.params = const_cast<double*>(d_params);
    d.state = d_state;
    d.out = d_mean;
    d.newton_fail = (unsigned long long*)(base + OFF_COUNTER);
    d.dreq = (DevRequest*)base;
    const DevRequest dr = make_dev_request(hr, d);
    k_noalloc_setup<<<dim3(1), dim3(WSUM_THREADS), 0, s>>>(dr, base, d_params + (uint64_t)P_MU * hr.N, hr.N, hr.dout);
    SP_HIP(hipGetLastError());
    const uint32_t block = MAX_BLOCK, waves_per_block = MAX_BLOCK / 64;
    const dim3 grid((hr.n_tiles + waves_per_block - 1) / waves_per_block);
    for (uint64_t sb = 0; sb < hr.total_steps || sb == 0; ) {
        const uint64_t se = std::min(sb + C, hr.total_steps);
        if (!launch_plain(hr.kind, grid, dim3(block), s, d.dreq, sb, se, partial, hr.n_tiles))
            return set_error(SP_ERR_UNSUPPORTED, "no kernel for this problem/solver/anisotropy/precision");
        SP_HIP(hipGetLastError());
        const uint64_t first_window = sb == 0 ? 0 : sb / OS + 1;
        const uint64_t rows = ((se - sb) / OS + (sb == 0 ? 1 : 0)) * hr.dout;
        const ReduceWork rw{partial, hr.n_tiles, 0u, (uint32_t)rows, (uint32_t)(first_window * hr.dout), d_mean, nullptr};
        k_noalloc_reduce<<<dim3(reduce_grid_x((uint32_t)rows), 1), dim3(256), 0, s>>>(rw);
        SP_HIP(hipGetLastError());
        if (se >= hr.total_steps) break;
        sb = se;
    }
    const uint64_t m = hr.n_out * hr.dout;
    k_noalloc_normalize<<<dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s>>>(d_mean, m, hr.dout, wsum);
    SP_HIP(hipGetLastError());
    return SP_OK;
}
