// float64 instantiation of the fused 2-D kernels (compiled without FP contraction).
#define HIS_REAL double
#include "swe2d.inc"
