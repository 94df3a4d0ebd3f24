// k_hop_f16.hip -- F16 instantiations of the hopping kernel (k_hop.inc).
#include "k_hop.inc"

namespace qcu {
void launch_hop_f16(int recon, int epi, const HopArgs& a, int block, int rb, hipStream_t st) {
  launch_hop_prec<F16>(recon, epi, a, block, rb, st);
}
}  // namespace qcu
