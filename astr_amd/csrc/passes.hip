// passes.hip -- the pass kernels of one launch shape (single case and
// batched), in register budgets MINB = 1 and 2.  Built once per shape with ASTR_SHAPE_ID = 0..12 so the
// shapes compile in parallel; flow.hip dispatches to launch_pass_s<ID>.
#include "flow_dev.hpp"

#ifndef ASTR_SHAPE_ID
#error "build with -DASTR_SHAPE_ID=<0..12>"
#endif

namespace astr {

#if ASTR_SHAPE_ID == 0
#define ASTR_LAUNCH launch_pass_s0
#define ASTR_BATCH launch_batch_s0
using ShapeCfg = CfgA;
#elif ASTR_SHAPE_ID == 1
#define ASTR_LAUNCH launch_pass_s1
#define ASTR_BATCH launch_batch_s1
using ShapeCfg = CfgB;
#elif ASTR_SHAPE_ID == 2
#define ASTR_LAUNCH launch_pass_s2
#define ASTR_BATCH launch_batch_s2
using ShapeCfg = CfgC;
#elif ASTR_SHAPE_ID == 3
#define ASTR_LAUNCH launch_pass_s3
#define ASTR_BATCH launch_batch_s3
using ShapeCfg = CfgD;
#elif ASTR_SHAPE_ID == 4
#define ASTR_LAUNCH launch_pass_s4
#define ASTR_BATCH launch_batch_s4
using ShapeCfg = CfgE;
#elif ASTR_SHAPE_ID == 5
#define ASTR_LAUNCH launch_pass_s5
#define ASTR_BATCH launch_batch_s5
using ShapeCfg = CfgF;
#elif ASTR_SHAPE_ID == 6
#define ASTR_LAUNCH launch_pass_s6
#define ASTR_BATCH launch_batch_s6
using ShapeCfg = CfgG;
#elif ASTR_SHAPE_ID == 7
#define ASTR_LAUNCH launch_pass_s7
#define ASTR_BATCH launch_batch_s7
using ShapeCfg = CfgH;
#elif ASTR_SHAPE_ID == 8
#define ASTR_LAUNCH launch_pass_s8
#define ASTR_BATCH launch_batch_s8
using ShapeCfg = CfgI;
#elif ASTR_SHAPE_ID == 9
#define ASTR_LAUNCH launch_pass_s9
#define ASTR_BATCH launch_batch_s9
using ShapeCfg = CfgJ;
#elif ASTR_SHAPE_ID == 10
#define ASTR_LAUNCH launch_pass_s10
#define ASTR_BATCH launch_batch_s10
using ShapeCfg = CfgK;
#elif ASTR_SHAPE_ID == 11
#define ASTR_LAUNCH launch_pass_s11
#define ASTR_BATCH launch_batch_s11
using ShapeCfg = CfgL;
#else
#define ASTR_LAUNCH launch_pass_s12
#define ASTR_BATCH launch_batch_s12
using ShapeCfg = CfgM;
#endif

hipError_t ASTR_LAUNCH(int prog, int minb, const FlowArgs& a, int dir, hipStream_t st) {
  return minb >= 2 ? launch_pass_c<ShapeCfg, 2>(prog, a, dir, st)
                   : launch_pass_c<ShapeCfg, 1>(prog, a, dir, st);
}

hipError_t ASTR_BATCH(int prog, int minb, const BatchLaunch& bl, int dir, const StageArgs& sa,
                      int nblocks, hipStream_t st) {
  return minb >= 2 ? launch_batch_c<ShapeCfg, 2>(prog, bl, dir, sa, nblocks, st)
                   : launch_batch_c<ShapeCfg, 1>(prog, bl, dir, sa, nblocks, st);
}

}  // namespace astr
