// ig_jk_lane64.hip -- FP64 lane-per-quartet J/K kernels (classes up to d)
// and the host helpers of their variants.
#define IG_LANE_T double
#define IG_LANE_FN launch_jk_lane
#define IG_LANE_HOST_HELPERS 1
#define IG_LANE_BATCH_FN launch_jk_batch
#define IG_LANE_CLASSES IG_FOR_CLASSES
#define IG_LANE_NEXT launch_jk_lane_fcls
#define IG_LANE_BATCH_NEXT launch_jk_batch_fcls
#include "ig_lane.h"
