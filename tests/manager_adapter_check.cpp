// Compiles include/stophys_gpu_simulation_manager.hpp -- the SimulationManager<prec, InstructionSet::GFX950>
// specialisation meant for the reference's build -- against a test double of the three names of the reference it
// uses (the InstructionSet enumeration with GFX950 added, the primary template SimulationManager<prec, set> and
// SerAr's file archive), and prints the request the constructed manager holds.  The settings object of the double
// carries the text of a parameter file and its archive writes that text out, which is what the reference's archive
// does with a SimulationManagerSettings it serialises (InputParams.cpp:224-226).
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <string>

namespace MyCEL::SystemInfo {
enum class InstructionSet { NONE, AVX, AVX2, AVX512, GFX950 };
}
namespace SimulationApplication {
template <typename prec, MyCEL::SystemInfo::InstructionSet set>
class SimulationManager;
}
namespace Settings {
template <typename prec>
struct SimulationManagerSettings {
    std::string text;
};
}  // namespace Settings
namespace SerAr {
enum class ArchiveOutputMode { Overwrite };
struct AllFileOutputArchiveWrapper {
    std::ofstream out;
    AllFileOutputArchiveWrapper(const std::filesystem::path& p, ArchiveOutputMode) : out(p) {}
    template <class T>
    void operator()(const T& settings) { out << settings.text; }
};
}  // namespace SerAr

#include "stophys_gpu_simulation_manager.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::ifstream in(argv[1]);
    const Settings::SimulationManagerSettings<double> settings{std::string(std::istreambuf_iterator<char>(in), {})};
    try {
        using Application = SimulationApplication::SimulationManager<double, MyCEL::SystemInfo::InstructionSet::GFX950>;
        Application manager{settings};
        const sp_request& r = manager.request();
        std::printf("%d %d %d %d %d %llu %llu %llu %llu %.17g %u %s\n", r.problem, r.solver, r.field, r.anisotropy, r.precision,
                    (unsigned long long)r.n_particles, (unsigned long long)r.n_steps, (unsigned long long)r.oversampling,
                    (unsigned long long)r.seed, r.dt, r.flags, manager.result_file().c_str());
        Application::ProgressFactor() = 1.0;
        return 0;
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
}
