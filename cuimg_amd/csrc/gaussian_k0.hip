// Standalone separable-Gaussian kernel instances for half widths 1, 5, 9, 13.
#include "gaussian_kernel_dev.hpp"

namespace cuimg {
CUIMG_GAUSS_INSTANCES(1)
CUIMG_GAUSS_INSTANCES(5)
CUIMG_GAUSS_INSTANCES(9)
CUIMG_GAUSS_INSTANCES(13)
}  // namespace cuimg
