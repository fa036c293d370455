// muca_team_r7_f64.hip -- instantiations of the team update kernel for Philox4x32-7
// with FP64 acceptance (both keyings); the dispatch is in muca_team.hip.
#include "team_run.h"

namespace muca {

template hipError_t launch_team_r<7, true, 0>(hipStream_t, const McsArgs &, const TeamLayout &);
template hipError_t launch_team_r<7, true, 1>(hipStream_t, const McsArgs &, const TeamLayout &);

}  // namespace muca
