// Host-buffer multiply. First form: whole-operand copies on the compute
// stream around the device-resident multiply.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "pm_state.h"

namespace pm {

namespace {
void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

void State::pipeline_init() {}
void State::pipeline_free() {}

void State::run_host(const uint32_t* a, const uint32_t* b, uint32_t* c) {
    const size_t bytes = coeff_words * 4;
    hip_check(hipMemcpyAsync(d_in[0], a, bytes, hipMemcpyHostToDevice, stream), "H2D a");
    hip_check(hipMemcpyAsync(d_in[1], b, bytes, hipMemcpyHostToDevice, stream), "H2D b");
    hip_check(enqueue_multiply(d_in[0], d_in[1], d_out, stream), "multiply");
    hip_check(hipMemcpyAsync(c, d_out, bytes, hipMemcpyDeviceToHost, stream), "D2H c");
    hip_check(hipStreamSynchronize(stream), "synchronize");
}

}  // namespace pm
