// Standalone separable-Gaussian kernel instances for half widths 3, 7, 11, 15.
#include "gaussian_kernel_dev.hpp"

namespace cuimg {
CUIMG_GAUSS_INSTANCES(3)
CUIMG_GAUSS_INSTANCES(7)
CUIMG_GAUSS_INSTANCES(11)
CUIMG_GAUSS_INSTANCES(15)
}  // namespace cuimg
