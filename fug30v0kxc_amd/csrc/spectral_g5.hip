// spectral_g5.hip -- transform kernels instantiated for one group of
// lengths (split so the kernels compile in parallel).
#include "spectral_impl.hpp"

#define SIZES_G5(X) X(768) X(1024)
MHD_SPECTRAL_GROUP(launch_spectral_g5, SIZES_G5)
