// k_hop.hip -- dispatch of the hopping kernel by precision and right-hand sides per
// thread (instantiations live in k_hop_<prec>_<rpt>.hip).
#include "qcu_kernels.h"

namespace qcu {
#define QCU_DECL(PN)                                                                         \
  void launch_hop_##PN##_1(int, int, const HopArgs&, int, int, hipStream_t);                 \
  void launch_hop_##PN##_2(int, int, const HopArgs&, int, int, hipStream_t);                 \
  void launch_hop_##PN##_3(int, int, const HopArgs&, int, int, hipStream_t);                 \
  void launch_hop_##PN##_4(int, int, const HopArgs&, int, int, hipStream_t);
QCU_DECL(f64)
QCU_DECL(f32)
QCU_DECL(f16)
#undef QCU_DECL

void launch_hop(int prec, int recon, int epi, const HopArgs& a, int block, int rb, int rpt, hipStream_t st) {
  typedef void (*Fn)(int, int, const HopArgs&, int, int, hipStream_t);
  static const Fn table[3][4] = {
      {launch_hop_f64_1, launch_hop_f64_2, launch_hop_f64_3, launch_hop_f64_4},
      {launch_hop_f32_1, launch_hop_f32_2, launch_hop_f32_3, launch_hop_f32_4},
      {launch_hop_f16_1, launch_hop_f16_2, launch_hop_f16_3, launch_hop_f16_4}};
  if (rpt < 1) rpt = 1;
  if (rpt > 4) rpt = 4;
  table[prec][rpt - 1](recon, epi, a, block, rb, st);
}
}  // namespace qcu
