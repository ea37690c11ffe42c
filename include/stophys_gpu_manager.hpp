// stophys_gpu_manager.hpp -- C++ SimulationManager surface over the C ABI.
//
// Header-only front-end with the public surface of the reference's
// SimulationApplication::SimulationManager<prec, InstructionSet>
// (Particle_Simulation/Simulator/SimulationManager.h:545-615): constructor from
// a parameter file, StartSimulationManager, waitUntilFinsihed (sic), isFinished,
// abort and the static progress accessors (updated per chunk while waiting).
// stophys_gpu_simulation_manager.hpp builds the reference's own
// SimulationManager<prec, InstructionSet::GFX950> on it.  The whole ensemble runs on one GPU
// through libstophys_gpu.so, its particles drawn on the device by the provider
// kernel (no per-particle array on the host); results are written like the
// reference's Mean_Results when the manager finishes.  Errors are rethrown as
// std::runtime_error, as the reference does.
#pragma once
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "stophys_gpu.h"

namespace SimulationApplication {

class SimulationManagerGFX950 {
public:
    using Parameters = std::string;   // path of a SimulationManagerSettings parameter file

    // the parameter file is read here; the device context is opened by StartSimulationManager
    explicit SimulationManagerGFX950(const Parameters& parameter_file, int device = 0, std::string result_file = {})
        : device_(device), result_file_(std::move(result_file)) {
        check(sp_settings_load(parameter_file.c_str(), &settings_));
        check(sp_settings_request(settings_, &req_, nullptr));
        if (result_file_.empty()) {
            const char* s = sp_settings_savefile(settings_);
            result_file_ = s && *s ? std::string(s) : std::string("Mean_Results");
            const auto dot = result_file_.find_last_of('.');
            if (dot != std::string::npos) result_file_ = result_file_.substr(0, dot);
            result_file_ += ".json";
        }
    }
    SimulationManagerGFX950(const SimulationManagerGFX950&) = delete;
    SimulationManagerGFX950& operator=(const SimulationManagerGFX950&) = delete;
    ~SimulationManagerGFX950() {
        if (ctx_) sp_ctx_destroy(ctx_);
        if (settings_) sp_settings_free(settings_);
    }

    // SimulationManager.h:549-574: starts the ensemble (asynchronously here)
    void StartSimulationManager() {
        if (started_) return;
        if (!ctx_) check(sp_ctx_create(device_, &ctx_));
        check(sp_settings_submit(ctx_, settings_, &id_));
        started_ = true;
    }
    // SimulationManager.h:577-584; reports progress while it waits
    void waitUntilFinsihed() {
        if (!started_) throw std::runtime_error("SimulationManager: not started");
        if (collected_) return;
        for (;;) {
            const int rc = sp_wait(ctx_, id_, 200);
            if (rc == SP_OK) break;
            if (rc != SP_ERR_TIMEOUT) check(rc);
            update_progress();
        }
        collect();
    }
    // SimulationManager.h:587-592
    bool isFinished() {
        if (!started_) return false;
        if (collected_) return true;
        if (sp_wait(ctx_, id_, 0) != SP_OK) {
            update_progress();
            return false;
        }
        collect();
        return true;
    }
    // SimulationManager.h:595-599
    void abort() {
        if (started_ && !collected_) sp_cancel(ctx_, id_);
    }
    // SimulationManager.h:601-615
    static double& ProgressModifier() { static double v{0.0}; return v; }
    static double& ProgressFactor() { static double v{1.0}; return v; }
    static std::atomic<std::size_t>& ProgressCache() { static std::atomic<std::size_t> v{0}; return v; }

    const sp_request& request() const { return req_; }    // the run the parameter file describes
    const sp_result& result() const { return res_; }
    const std::vector<double>& time() const { return time_; }
    const std::vector<double>& mean() const { return mean_; }
    const std::string& result_file() const { return result_file_; }
    const std::string& single_result_file() const { return single_file_; }   // empty unless singles were saved
    const std::vector<double>& trajectories() const { return traj_; }        // [n_out][dim_out][n_saved]

private:
    static void check(int rc) {
        if (rc != SP_OK) throw std::runtime_error(std::string("stophys_gpu: ") + sp_last_error());
    }
    // SimulationManager.h:177-214 with the fraction of steps done in place of the fraction of finished
    // simulations: the percentage (scaled by ProgressFactor, shifted by ProgressModifier) goes to ProgressCache
    // and, when it changes, a 70-character bar to stderr
    void update_progress() {
        double f = 0.0;
        if (sp_progress(ctx_, id_, &f) != SP_OK) return;
        const auto progress = (std::size_t)std::llround(f * 100.0 * ProgressFactor() + ProgressModifier() * 100.0);
        if (ProgressCache().exchange(progress) == progress) return;
        constexpr int barlength = 70;
        std::string bar(1, '[');
        for (std::size_t i = 0; i < barlength; ++i) bar += i * 100 < progress * barlength ? '=' : ' ';
        std::fprintf(stderr, "%s] %zu%%\n", bar.c_str(), progress);
    }
    void collect() {
        uint64_t n_out = 0;
        uint32_t dy = 0, dw = 0, dout = 0;
        check(sp_request_shape(&req_, &n_out, &dy, &dw, &dout));
        time_.assign(n_out, 0.0);
        field_.assign(n_out * 3, 0.0);
        mean_.assign(n_out * dout, 0.0);
        weight_.assign(dout, 0.0);
        res_ = sp_result{};
        res_.time = time_.data();
        res_.field = field_.data();
        res_.mean = mean_.data();
        res_.weight = weight_.data();
        if (req_.save_interval) {   // Save_single_simulations: every Save_intervall-th particle's own trajectory
            traj_.assign(n_out * dout * (req_.n_particles / req_.save_interval), 0.0);
            res_.trajectories = traj_.data();
        }
        check(sp_result_get(ctx_, id_, &res_));
        check(sp_settings_write_result(settings_, &res_, result_file_.c_str()));
        if (res_.n_saved) {
            const char* sf = sp_settings_single_savefile(settings_);
            std::string path = sf && *sf ? std::string(sf) : std::string();
            if (path.empty() || path == result_file_) {   // the reference refuses one file for both
                path = result_file_;
                const auto dot = path.find_last_of('.');
                path = (dot == std::string::npos ? path : path.substr(0, dot)) + "_single.json";
            }
            check(sp_settings_write_single_results(settings_, &res_, path.c_str()));
            single_file_ = path;
        }
        ProgressCache().store((std::size_t)std::llround(100.0 * ProgressFactor() + ProgressModifier() * 100.0));
        collected_ = true;
    }

    int device_ = 0;
    sp_settings* settings_ = nullptr;
    sp_ctx* ctx_ = nullptr;
    sp_request req_{};
    uint64_t id_ = 0;
    bool started_ = false, collected_ = false;
    std::string result_file_, single_file_;
    sp_result res_{};
    std::vector<double> time_, field_, mean_, weight_, traj_;
};

}  // namespace SimulationApplication
