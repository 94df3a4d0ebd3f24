/*
 * qcu_shift_storage.h -- the reference's storage shifts between its AoS interface
 * layouts and its coalesced structure-of-arrays layouts, as C entry points of
 * libqcu.so.
 *
 * Same names, enums, argument order and meaning as the reference header
 * include/qcu_shift_storage.cuh:3-16 (src/qcu_shift_storage.cu:74-194).  Pointers are
 * HIP device pointers; Lx is the full x extent (even); Vh = Lx*Ly*Lz*Lt/2.
 *
 *   vector  NON_COALESCED  complex128 [Vh][4][3]        one half spinor
 *           COALESCED      double     [24][Vh]          plane i = real i of the 24 of a site
 *   gauge   NON_COALESCED  complex128 [4][2][Vh][3][3]
 *           COALESCED      double     [4][2][12][Vh]    rows 0 and 1 of every link
 *
 * The gauge shift back to NON_COALESCED, which the reference leaves empty, rebuilds
 * row 2 = conj(row0 x row1) (src/qcu_wilson_dslash_neo.cu:73-78).
 *
 * Like the reference, both functions synchronise before returning.  Errors follow
 * qcu.h: a message on stderr, the error recorded for qcuGetLastError(), and exit(-1)
 * unless QCU_EXIT_ON_ERROR=0.  dst and src must not overlap.
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

/* include/qcu_shift_storage.cuh:3-6 */
enum MemoryStorage {
  NON_COALESCED = 0,
  COALESCED = 1,
};

/* include/qcu_shift_storage.cuh:8-11 */
enum ShiftDirection {
  TO_COALESCE = 0,
  TO_NON_COALESCE = 1,
};

/* Replaces shiftVectorStorage, include/qcu_shift_storage.cuh:14 (src/qcu_shift_storage.cu:166-180). */
void shiftVectorStorage(void* dst_vec, void* src_vec, int shift_direction, int Lx, int Ly, int Lz, int Lt);

/* Replaces shiftGaugeStorage, include/qcu_shift_storage.cuh:16 (src/qcu_shift_storage.cu:183-194). */
void shiftGaugeStorage(void* dst_vec, void* src_vec, int shift_direction, int Lx, int Ly, int Lz, int Lt);

#ifdef __cplusplus
}
#endif
