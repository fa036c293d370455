// muca_team_r10_f64.hip -- instantiations of the team update kernel for Philox4x32-10
// with FP64 acceptance (both keyings); the dispatch is in muca_team.hip.
#include "team_run.h"

namespace muca {

template hipError_t launch_team_r<10, true, 0>(hipStream_t, const McsArgs &, const TeamLayout &);
template hipError_t launch_team_r<10, true, 1>(hipStream_t, const McsArgs &, const TeamLayout &);

}  // namespace muca
