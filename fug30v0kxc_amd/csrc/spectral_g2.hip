// spectral_g2.hip -- transform kernels instantiated for one group of
// lengths (split so the kernels compile in parallel).
#include "spectral_impl.hpp"

#define SIZES_G2(X) X(160) X(192) X(240) X(256)
MHD_SPECTRAL_GROUP(launch_spectral_g2, SIZES_G2)
