// Standalone separable-Gaussian kernel instances for half widths 2, 6, 10, 14.
#include "gaussian_kernel_dev.hpp"

namespace cuimg {
CUIMG_GAUSS_INSTANCES(2)
CUIMG_GAUSS_INSTANCES(6)
CUIMG_GAUSS_INSTANCES(10)
CUIMG_GAUSS_INSTANCES(14)
}  // namespace cuimg
