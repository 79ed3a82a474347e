// FP32 batched TSQR kernels (see batch.hpp).
#include "batch_launch_impl.hpp"

namespace tsqr {
template int batch_launch<float>(const BatchDev&, void*, hipStream_t);
}
