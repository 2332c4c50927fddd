// ig_jk_lane32.hip -- lane-per-quartet J/K kernels with FP32 tables and
// block (mixed-precision quartets; roots, weights and prefactors in FP64, J/K
// accumulated in FP64).
#define IG_LANE_T float
#define IG_LANE_FN launch_jk_lane_f32
#include "ig_lane.h"
