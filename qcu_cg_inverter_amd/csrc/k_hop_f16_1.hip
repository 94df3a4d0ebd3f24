// k_hop_f16_1.hip -- F16 hopping kernels with 1 right-hand side(s) per thread (k_hop.inc).
#include "k_hop.inc"
QCU_HOP_INSTANTIATE(f16, F16, 1)
