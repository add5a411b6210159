// g2p2g.hip — the plain fused substep kernel k_g2p2g (G2P of substep s fused with P2G of s+1, analytic
// closure) and its launcher, in a translation unit of their own: the Makefile compiles this file with LLVM's
// max-ilp machine scheduler (-mllvm -amdgpu-sched-strategy=max-ilp), which gives k_g2p2g a schedule about
// 1 % faster on the 4,194,304-particle mixed scene (62.0-62.3 -> 61.5-61.6 us per launch, three pairs) and
// the same speed on the 64-scene batch.  The learned kernels keep the default scheduler: under max-ilp the
// 16x16 learned kernel drops to a schedule that serialises its SiLU table reads and runs 9 % slower.  The
// code itself is in kernels.hip.
#define MPM_PLAIN_FUSED_TU 1
#include "kernels.hip"
