// k_hop_f32_3.hip -- F32 hopping kernels with 3 right-hand side(s) per thread (k_hop.inc).
#include "k_hop.inc"
QCU_HOP_INSTANTIATE(f32, F32, 3)
