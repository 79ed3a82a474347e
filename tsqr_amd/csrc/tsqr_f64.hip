// FP64 TSQR kernels (see tsqr_kernels.hpp).
#include "tsqr_launch_impl.hpp"

namespace tsqr {
template int tsqr_launch<double>(const TreeDesc&, double*, int, double*, int, void*, const LaunchOpts&, hipStream_t);
}
