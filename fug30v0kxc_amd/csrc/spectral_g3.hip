// spectral_g3.hip -- transform kernels instantiated for one group of
// lengths (split so the kernels compile in parallel).
#include "spectral_impl.hpp"

#define SIZES_G3(X) X(320) X(384) X(480)
MHD_SPECTRAL_GROUP(launch_spectral_g3, SIZES_G3)
