// float32-bulk instantiation of the fused 2-D kernels.
#define HIS_REAL float
#include "swe2d.inc"
