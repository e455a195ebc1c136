// spectral_g0.hip -- transform kernels instantiated for one group of
// lengths (split so the kernels compile in parallel).
#include "spectral_impl.hpp"

#define SIZES_G0(X) X(8) X(12) X(16) X(20) X(24) X(30) X(32) X(40) X(48) X(60)
MHD_SPECTRAL_GROUP(launch_spectral_g0, SIZES_G0)
