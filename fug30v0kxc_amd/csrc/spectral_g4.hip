// spectral_g4.hip -- transform kernels instantiated for one group of
// lengths (split so the kernels compile in parallel).
#include "spectral_impl.hpp"

#define SIZES_G4(X) X(512) X(640)
MHD_SPECTRAL_GROUP(launch_spectral_g4, SIZES_G4)
