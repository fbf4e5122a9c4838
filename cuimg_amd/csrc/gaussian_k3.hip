// Standalone separable-Gaussian kernel instances for half widths 4, 8, 12, 16.
#include "gaussian_kernel_dev.hpp"

namespace cuimg {
CUIMG_GAUSS_INSTANCES(4)
CUIMG_GAUSS_INSTANCES(8)
CUIMG_GAUSS_INSTANCES(12)
CUIMG_GAUSS_INSTANCES(16)
}  // namespace cuimg
