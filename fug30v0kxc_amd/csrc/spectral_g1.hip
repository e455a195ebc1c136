// spectral_g1.hip -- transform kernels instantiated for one group of
// lengths (split so the kernels compile in parallel).
#include "spectral_impl.hpp"

#define SIZES_G1(X) X(64) X(80) X(96) X(120) X(128)
MHD_SPECTRAL_GROUP(launch_spectral_g1, SIZES_G1)
