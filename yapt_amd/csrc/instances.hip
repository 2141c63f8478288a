// Template kernel instances of libyapt_pt, one group per compilation (-DYPT_INST_GROUP=n, n in
// [0, YPT_INST_GROUPS)), so the device code of the library compiles in parallel. The host
// translation unit (ypt_lib.hip) declares every instance extern and launches them.
#define YPT_INSTANCE_TU 1
#include "kernels/instances.h"

#ifndef YPT_INST_GROUP
#error "YPT_INST_GROUP selects the group of instances this compilation defines"
#endif

namespace ypt {
#if YPT_INST_GROUP == 0
YPT_INST_TRAV(template, 2)
#elif YPT_INST_GROUP == 1
YPT_INST_TRAV(template, 8)
#elif YPT_INST_GROUP == 2
YPT_INST_SHADE(template, true)
#elif YPT_INST_GROUP == 3
YPT_INST_SHADE(template, false)
#elif YPT_INST_GROUP == 4
YPT_INST_BDPT_LIGHT(template, 2)
YPT_INST_BDPT_LIGHT(template, 8)
#elif YPT_INST_GROUP == 5
YPT_INST_BDPT_CAMERA(template, 2, false)
#elif YPT_INST_GROUP == 6
YPT_INST_BDPT_CAMERA(template, 2, true)
#elif YPT_INST_GROUP == 7
YPT_INST_BDPT_CAMERA(template, 8, false)
#elif YPT_INST_GROUP == 8
YPT_INST_BDPT_CAMERA(template, 8, true)
#elif YPT_INST_GROUP == 9
YPT_INST_BDPT_CSHADE(template)
#elif YPT_INST_GROUP == 10
YPT_INST_FUSED(template)
#else
#error "no such instance group"
#endif
}  // namespace ypt
