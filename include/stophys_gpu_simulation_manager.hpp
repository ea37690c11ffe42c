// stophys_gpu_simulation_manager.hpp -- SimulationManager<prec, InstructionSet::GFX950> for the reference's app.
//
// Included in the reference's build after Simulator/SimulationManager.h (whose primary template
// SimulationApplication::SimulationManager<prec, MyCEL::SystemInfo::InstructionSet set>, SimulationManager.h:76-77,
// it specialises) and after arch/InstructionSets.hpp:11 gains the enumerator GFX950, so that
// runAppWithArch<InstructionSet::GFX950> (app/StoPhysApp_MultiArch.cpp:27-38) runs unchanged:
//
//     Application SimManager{ std::move(input.getAppParams()) };    // Settings::SimulationManagerSettings<PREC>
//     SimManager.StartSimulationManager();
//     SimManager.waitUntilFinsihed();
//
// The constructor takes the reference's Settings::SimulationManagerSettings<prec> (SimulationManager.h:545-547)
// and writes it with the reference's own archive, as InputParams.cpp:224-226 writes its settings, to a temporary
// parameter file; the GPU backend's front-end reads that file (the reference's .ini schema, particle files relative
// to the working directory as ParticleProvider.h:275-277 opens them) and the ensemble runs on one GPU.
#pragma once
#include <filesystem>
#include <string>
#include <system_error>

#include <unistd.h>

#include "stophys_gpu_manager.hpp"

namespace SimulationApplication {

template <typename prec>
class SimulationManager<prec, MyCEL::SystemInfo::InstructionSet::GFX950> : public SimulationManagerGFX950 {
public:
    using Parameters = Settings::SimulationManagerSettings<prec>;

    explicit SimulationManager(const Parameters& params, int device = 0)
        : SimulationManagerGFX950(ParameterFile(params).path, device) {}

private:
    // the settings as a parameter file for the lifetime of the constructor call (the front-end reads it whole)
    struct ParameterFile {
        std::string path;
        explicit ParameterFile(const Parameters& params) {
            static std::atomic<unsigned> serial{0};
            const auto p = std::filesystem::temp_directory_path() /
                           ("stophys_gfx950_" + std::to_string(::getpid()) + "_" + std::to_string(serial++) + ".ini");
            {
                SerAr::AllFileOutputArchiveWrapper archive(p, SerAr::ArchiveOutputMode::Overwrite);
                archive(params);
            }
            path = p.string();
        }
        ~ParameterFile() {
            std::error_code ec;
            std::filesystem::remove(path, ec);
        }
    };
};

}  // namespace SimulationApplication
