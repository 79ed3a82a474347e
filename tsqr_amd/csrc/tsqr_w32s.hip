// TSQR launch sequence (tsqr_launch_impl.hpp) for float matrices of padded width 32.
#include "tsqr_launch_impl.hpp"

namespace tsqr {
TSQR_WIDTH_UNIT(float, 32);
}
