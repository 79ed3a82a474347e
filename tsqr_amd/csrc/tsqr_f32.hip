// FP32 TSQR kernels (see tsqr_kernels.hpp).
#include "tsqr_launch_impl.hpp"

namespace tsqr {
template int tsqr_launch<float>(const TreeDesc&, float*, int, float*, int, void*, const LaunchOpts&, hipStream_t);
}
