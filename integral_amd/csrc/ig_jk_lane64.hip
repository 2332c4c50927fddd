// ig_jk_lane64.hip -- FP64 lane-per-quartet J/K kernels and the host helpers
// of their variants.
#define IG_LANE_T double
#define IG_LANE_FN launch_jk_lane
#define IG_LANE_HOST_HELPERS 1
#include "ig_lane.h"
