// k_hop_f32_2.hip -- F32 hopping kernels with 2 right-hand side(s) per thread (k_hop.inc).
#include "k_hop.inc"
QCU_HOP_INSTANTIATE(f32, F32, 2)
