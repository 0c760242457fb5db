// mpm_kernels_s1.hip -- the step kernels for tile shape 1 (TileShape<DIM, 1>), in namespace mpmk_s1
#define MPM_TILE_SHAPE 1
#include "mpm_kernels.hip"
