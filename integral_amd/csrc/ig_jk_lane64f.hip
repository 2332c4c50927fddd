// ig_jk_lane64f.hip -- FP64 lane-per-quartet J/K kernels of the f classes
// with blocks of at most four register slices.
#define IG_LANE_T double
#define IG_LANE_FN launch_jk_lane_fcls
#define IG_LANE_BATCH_FN launch_jk_batch_fcls
#define IG_LANE_CLASSES IG_FOR_F_LANE_CLASSES
#include "ig_lane.h"
