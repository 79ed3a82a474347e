// FP64 TSQR kernels (see tsqr_kernels.hpp) and the device-side checker (check.hpp).
#include "check.hpp"
#include "tsqr_launch_impl.hpp"

namespace tsqr {
template int tsqr_launch<double>(const TreeDesc&, double*, int, double*, int, void*, const LaunchOpts&, hipStream_t);
template int check_launch<double>(int, int, const double*, int, const double*, int, const double*, int, void*, size_t*,
                                  hipStream_t);
}
