// sp_host.h -- internal host-side declarations of the engine.
#pragma once
#include <stdint.h>
#include <string>

#include "../../include/stophys_gpu.h"
#include "sp_types.h"

namespace sp {

int request_dims(int problem, uint32_t* dy, uint32_t* dw, uint32_t* dout);
void compute_params(const sp_request& rq, const sp_particles& pt, double* out, uint64_t stride);
void time_and_field(const sp_request& rq, double* times, double* fields);
// the request's field as FieldParams (angular frequencies, rectangular edge constants)
FieldParams field_params(const sp_request& rq);

// thread-local error message for sp_last_error()
int set_error(int code, const std::string& msg);

}  // namespace sp
