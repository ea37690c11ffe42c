// stophys_gpu_app.cpp -- command line front-end with the options of the reference's
// StoPhysApp_MultiArch (app/InputParams.cpp:46-76; app/StoPhysApp_MultiArch.cpp:27-93):
//   --parameter_file=FILE     the SimulationManagerSettings .ini file
//   --instruction_set=GFX950  the only value this program runs (the GPU backend)
//   --result_file=FILE        where Mean_Results are written (default: Savefile with .json)
//   --device=N                GPU index
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../../include/stophys_gpu_manager.hpp"

int main(int argc, char** argv) {
    std::string param, iset = "GFX950", result;
    int device = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&](const char* key) -> const char* {
            const size_t n = std::strlen(key);
            if (a.compare(0, n, key) == 0 && a.size() > n && a[n] == '=') return argv[i] + n + 1;
            if (a == key && i + 1 < argc) return argv[++i];
            return nullptr;
        };
        if (const char* v = val("--parameter_file")) param = v;
        else if (const char* v2 = val("--instruction_set")) iset = v2;
        else if (const char* v3 = val("--result_file")) result = v3;
        else if (const char* v4 = val("--device")) device = std::atoi(v4);
        else { std::fprintf(stderr, "unknown option %s\n", argv[i]); return 2; }
    }
    if (param.empty()) { std::fprintf(stderr, "usage: %s --parameter_file=FILE [--instruction_set=GFX950]\n", argv[0]); return 2; }
    if (iset != "GFX950") { std::fprintf(stderr, "instruction set %s is not built here (GFX950 only)\n", iset.c_str()); return 2; }
    try {   // StoPhysApp_MultiArch.cpp:78-90 catches and reports exceptions the same way
        const auto t0 = std::chrono::steady_clock::now();
        SimulationApplication::SimulationManagerGFX950 sim(param, device, result);
        sim.StartSimulationManager();
        sim.waitUntilFinsihed();
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const sp_result& r = sim.result();
        std::printf("Simulation Manager: finished %.0f particle-steps in %.3f s (device %.3f s, %.3g particle-steps/s); "
                    "results in %s\n", r.particle_steps, s, r.device_ms * 1e-3, r.particle_steps / s, sim.result_file().c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
    return 0;
}
