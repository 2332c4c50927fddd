// ig_jk_lane32.hip -- lane-per-quartet J/K kernels with FP32 tables and
// block (mixed-precision quartets; roots, weights and prefactors in FP64, J/K
// accumulated in FP64), classes up to d.
#define IG_LANE_T float
#define IG_LANE_FN launch_jk_lane_f32
#define IG_LANE_BATCH_FN launch_jk_batch_f32
#define IG_LANE_CLASSES IG_FOR_CLASSES
#define IG_LANE_NEXT launch_jk_lane_f32_fcls
#define IG_LANE_BATCH_NEXT launch_jk_batch_f32_fcls
#include "ig_lane.h"
