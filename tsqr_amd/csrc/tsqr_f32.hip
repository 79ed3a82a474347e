// FP32 TSQR kernels (see tsqr_kernels.hpp) and the device-side checker (check.hpp).
#include "check.hpp"
#include "tsqr_launch_impl.hpp"

namespace tsqr {
template int tsqr_launch<float>(const TreeDesc&, float*, int, float*, int, void*, const LaunchOpts&, hipStream_t);
template int check_launch<float>(int, int, const float*, int, const float*, int, const float*, int, void*, size_t*,
                                  hipStream_t);
}
