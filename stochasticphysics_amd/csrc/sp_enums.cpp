// sp_enums.cpp -- selector strings of the reference (CPU only).
// Settings/SolverSettings.cpp:10-20, Settings/ProblemSettings.h:63-70,
// Fields/FieldList.h:51-59, Problems/Anisotropy/AnisotropyList.h:48-54.
#include <cstring>
#include <string>

#include "sp_host.h"

namespace {
struct Entry { const char* kind; const char* name; int value; };
const Entry kEntries[] = {
    {"problem", "undefined", 0}, {"problem", "Neel", SP_PROBLEM_NEEL},
    {"problem", "NeelSpherical", SP_PROBLEM_NEEL_SPHERICAL}, {"problem", "BrownAndNeel", SP_PROBLEM_BROWN_AND_NEEL},
    {"problem", "BrownAndNeelEulerSpherical", 4},
    {"solver", "undefined", 0}, {"solver", "EulerMaruyama", SP_SOLVER_EULER_MARUYAMA},
    {"solver", "EulerMaruyama_Normalized", 2}, {"solver", "Implicit_Midpoint", SP_SOLVER_IMPLICIT_MIDPOINT},
    {"solver", "Implicit_Midpoint_GSL", 4}, {"solver", "Implicit_Midpoint_GSL_Derivative_Free", 5},
    {"solver", "EulerMaruyama_Normalized", SP_SOLVER_EULER_MARUYAMA_NORMALIZED},
    {"solver", "Millstein", SP_SOLVER_MILLSTEIN}, {"solver", "Heun_Strong", SP_SOLVER_HEUN_STRONG},
    {"solver", "Heun_NotConsistent", 8}, {"solver", "WeakTest", 9}, {"solver", "ExplicitStrong_1.0", 10},
    {"field", "none", SP_FIELD_ZERO}, {"field", "constant", SP_FIELD_CONSTANT}, {"field", "sinusoidal", SP_FIELD_SINUSOIDAL},
    {"field", "lissajous", SP_FIELD_LISSAJOUS}, {"field", "triangular", SP_FIELD_TRIANGULAR}, {"field", "rectangular", SP_FIELD_RECTANGULAR},
    {"field", "sinc", SP_FIELD_SINC},
    {"anisotropy", "undefined", 0}, {"anisotropy", "uniaxial", SP_ANISO_UNIAXIAL}, {"anisotropy", "cubic", SP_ANISO_CUBIC},
    {"anisotropy", "mixed", 3}, {"anisotropy", "uniaxialcubic", 4},
};
}  // namespace

extern "C" {
int sp_parse_enum(const char* kind, const char* name, int32_t* value) {
    if (!kind || !name || !value) return sp::set_error(SP_ERR_INVALID_ARG, "NULL argument");
    for (const Entry& e : kEntries)
        if (!std::strcmp(e.kind, kind) && !std::strcmp(e.name, name)) { *value = e.value; return SP_OK; }
    // the reference's from_string throws on unknown names (e.g. Settings/SolverSettings.cpp:39-47)
    return sp::set_error(SP_ERR_INVALID_ARG, std::string("unknown ") + kind + " '" + name + "'");
}
const char* sp_enum_name(const char* kind, int32_t value) {
    if (!kind) return nullptr;
    for (const Entry& e : kEntries)
        if (!std::strcmp(e.kind, kind) && e.value == value) return e.name;
    return nullptr;
}
}
