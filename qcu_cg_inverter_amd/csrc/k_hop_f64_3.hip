// k_hop_f64_3.hip -- F64 hopping kernels with 3 right-hand side(s) per thread (k_hop.inc).
#include "k_hop.inc"
QCU_HOP_INSTANTIATE(f64, F64, 3)
