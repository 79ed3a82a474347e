// FP64 batched TSQR kernels (see batch.hpp).
#include "batch_launch_impl.hpp"

namespace tsqr {
template int batch_launch<double>(const BatchDev&, void*, hipStream_t);
}
