// sp_types.h -- request and work descriptors shared by the host runtime and the kernels.
#pragma once
#include <stdint.h>

#include "sp_field.h"

namespace sp {

// largest work-group of the simulation kernels (their __launch_bounds__)
constexpr uint32_t MAX_BLOCK = 256;

// Per-particle parameter block, param-major in HBM: params[k * N + p].  Filled on
// the host by the reference's parameter calculators (sp_params.cpp).
enum ParamIndex : int {
    P_A = 0,      // NeelFactor1  a = -gamma/(1+alpha^2)              ParameterCalculatorNeel.h:67
    P_B = 1,      // NeelFactor2  b = |a| alpha                        :68
    P_SIG = 2,    // NoisePrefactor sigma = sqrt(2 alpha kB T/(gamma mu)) :80-86
    P_D = 3,      // DriftPrefactor D = -a^2 (1+alpha^2) sigma^2      :77
    P_MU = 4,     // saturation moment mu = Ms V (ensemble weight)    MagneticProperties.h:166
    P_K1 = 5,     // field prefactor -2K/Ms (UniaxialAnisotropy.h:71-77, CubicAnisotropy.h:93)
    P_PT = 6,     // torque prefactor -2K V (UniaxialAnisotropy.h:79-85), BrownAndNeel only
    P_C = 7,      // BrownPrefactor c = 1/(6 eta V_h)                 ParameterCalculatorBrown.h:43
    P_CF = 8,     // Brown_F_Noise cF = c sqrt(2 kB T 6 eta V_h)     :45-56
    P_DD = 9,     // d = c mu                                         BrownAndNeelRelaxation.inl:48
    P_AX = 10,    // body axes x(3) y(3) z(3): params 10..18 (Neel, NeelSpherical)
    P_KU = 19,    // uniaxialcubic: field prefactor -2K_u/Ms of the uniaxial part (P_K1 holds -2K_c/Ms)
    P_U = 20,     // uniaxialcubic: easy axis R^T z of the uniaxial part (3), UniaxialCubicAnisotropy.hpp:114;
                  // BrownAndNeelEulerSpherical: -2K_u V, and R itself in the axis rows (the body axes are state)
    NPARAM = 23
};

// Uniform (per request) description of one ensemble on the device.
struct DevRequest {
    double dt, sqdt;
    double tmin, tmax;             // NeelSpherical chart rotation window (rotation when theta outside)
    double ctmin, ctmax;           // cos(tmin), cos(tmax) (a state in the rotated chart is tested through cos theta)
    double bmin, bmax;             // BrownAndNeelEulerSpherical: Euler chart window of the body angle theta
    double bcmin, bcmax;           // cos(bmin), cos(bmax) (a body in the rotated chart is tested through cos theta)
    double newton_tol;
    FieldParams field;             // applied field (sp_field.h)
    uint64_t n_particles;
    uint64_t seed;
    uint64_t replay_count;         // normals available per particle in replay mode
    uint32_t oversampling;
    uint32_t flags;                // RQ_BN_SIMPLE, RQ_REPLAY, RQ_NORMALIZE
    uint32_t newton_max_iter;
    uint32_t n_tiles;              // ceil(N / 64)
    uint32_t kind;                 // kernel class (see kernel_kind)
    const double* params;          // [NPARAM][N]
    double* state;                 // [DY][N]
    const double* replay;          // [K][N] standard normals (replay mode)
    unsigned long long* newton_fail;
    double* traj;                  // single-simulation outputs [window][DOUT][n_saved], or null
    uint64_t n_saved;              // particles p with (p + 1) % save_interval == 0
    uint32_t save_interval;
    uint32_t pad;
};

enum : uint32_t { RQ_BN_SIMPLE = 1u, RQ_REPLAY = 2u, RQ_NORMALIZE = 4u };

// One unit of work for one wave: 64 particles of one request for a step range.
// Window rows: row r of the partial buffer holds window first_window + r;
// window 0 is the initial state, window w >= 1 averages steps ((w-1) OS, w OS].
struct TileWork {
    uint32_t req;              // index into the DevRequest table
    uint32_t tile;             // 64-particle tile index within the request
    uint64_t step_begin;       // steps [step_begin, step_end), multiples of OS; step_begin % 4 == 0
    uint64_t step_end;
    double* partial;           // partial[(row * DOUT + c) * pstride + pslot]
    uint32_t pstride;
    uint32_t pslot;
};

// One request's share of a persistent launch: tiles [tile_base, tile_base + n_tiles) of
// the launch's global tile range belong to request `req`, steps [step_begin, step_end).
struct SlotWork {
    uint64_t step_begin, step_end;
    double* partial;           // partial[(row * DOUT + c) * n_tiles + tile]
    uint32_t req;
    uint32_t tile_base;
    uint32_t n_tiles;
    uint32_t pad;
};

// kernel class id = problem * 128 + solver * 8 + anisotropy slot * 2 + mixed (host side bookkeeping);
// anisotropy slots: uniaxial 0, cubic 1, uniaxialcubic 2
inline constexpr uint32_t kernel_kind(int problem, int solver, int aniso, int mixed) {
    return (uint32_t)(problem * 128 + solver * 8 + (aniso == 4 ? 2 : aniso - 1) * 2 + mixed);
}
inline constexpr bool kind_is_uniaxialcubic(uint32_t kind) { return ((kind >> 1) & 3u) == 2u; }

// Sum of the ensemble weights (the mu of every particle, MeanSimulationResult.h:63-75) in one fixed order, the
// same on the host and on the device: WSUM_THREADS partial sums, partial t adding particles t, t + WSUM_THREADS,
// ... in turn, then folded by halves (partial t += partial t + h for t < h, h = WSUM_THREADS / 2, / 4, ..., 1).
inline constexpr uint32_t WSUM_THREADS = 256;
inline double weight_sum_fixed(const double* w, uint64_t n) {
    double part[WSUM_THREADS];
    for (uint32_t t = 0; t < WSUM_THREADS; ++t) {
        double s = 0.0;
        for (uint64_t p = t; p < n; p += WSUM_THREADS) s += w[p];
        part[t] = s;
    }
    for (uint32_t h = WSUM_THREADS / 2; h > 0; h >>= 1)
        for (uint32_t t = 0; t < h; ++t) part[t] += part[t + h];
    return part[0];
}

// blocks along x of a reduction launch over `rows` rows: the grid of a launch may hold at most 2^32 - 1
// work-items per dimension (16,777,215 blocks of 256), so a block reduces every gridDim.x-th row
inline constexpr uint32_t REDUCE_MAX_GRID_X = 65535;
inline constexpr uint32_t reduce_grid_x(uint32_t rows) { return rows < REDUCE_MAX_GRID_X ? rows : REDUCE_MAX_GRID_X; }

// deterministic reduction of partial sums (one block per output row, rows strided over the grid)
struct ReduceWork {
    const double* partial;     // rows of pstride doubles
    uint32_t pstride;
    uint32_t n_check;          // leading slots also copied one by one into check_out (mixed-precision check)
    uint32_t n_rows;           // rows (windows * DOUT)
    uint32_t row_base;         // first output row
    double* out;               // out[row_base + r]
    double* check_out;         // nullable; check_out[(row_base + r) * n_check + slot]
};

}  // namespace sp
