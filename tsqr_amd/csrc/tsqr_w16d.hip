// TSQR launch sequence (tsqr_launch_impl.hpp) for double matrices of padded width 16.
#include "tsqr_launch_impl.hpp"

namespace tsqr {
TSQR_WIDTH_UNIT(double, 16);
}
