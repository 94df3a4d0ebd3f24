// k_hop_f32_1.hip -- F32 hopping kernels with 1 right-hand side(s) per thread (k_hop.inc).
#include "k_hop.inc"
QCU_HOP_INSTANTIATE(f32, F32, 1)
