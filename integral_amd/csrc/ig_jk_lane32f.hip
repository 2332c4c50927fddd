// ig_jk_lane32f.hip -- lane-per-quartet J/K kernels of the f classes with
// FP32 tables and block (mixed-precision quartets).
#define IG_LANE_T float
#define IG_LANE_FN launch_jk_lane_f32_fcls
#define IG_LANE_BATCH_FN launch_jk_batch_f32_fcls
#define IG_LANE_CLASSES IG_FOR_F_LANE_CLASSES
#include "ig_lane.h"
