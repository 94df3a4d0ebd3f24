// k_hop_f64.hip -- F64 instantiations of the hopping kernel (k_hop.inc).
#include "k_hop.inc"

namespace qcu {
void launch_hop_f64(int recon, int epi, const HopArgs& a, int block, int rb, hipStream_t st) {
  launch_hop_prec<F64>(recon, epi, a, block, rb, st);
}
}  // namespace qcu

namespace qcu {
void launch_hop_f32(int recon, int epi, const HopArgs& a, int block, int rb, hipStream_t st);
void launch_hop_f16(int recon, int epi, const HopArgs& a, int block, int rb, hipStream_t st);

void launch_hop(int prec, int recon, int epi, const HopArgs& a, int block, int rb, hipStream_t st) {
  if (prec == F64) launch_hop_f64(recon, epi, a, block, rb, st);
  else if (prec == F32) launch_hop_f32(recon, epi, a, block, rb, st);
  else launch_hop_f16(recon, epi, a, block, rb, st);
}
}  // namespace qcu
