/*
 * qcu.h -- drop-in C entry points of the MI355X qcu library (libqcu.so).
 *
 * Same structs, symbols, argument meanings and layouts as the reference header
 * include/qcu.h:1-25 of YellowBicycleee/Qcu_Cg_Inverter (cg_inverter/).  All field
 * pointers are HIP device pointers to complex128 data in the reference layouts:
 *   gauge        [mu 4][parity 2][Vh][3][3]   U_mu(x) is the link x -> x+mu
 *   half spinor  [Vh][spin 4][colour 3]       half-site index ((t*Lz+z)*Ly+y)*(Lx/2)+xh
 *   full spinor  [parity 2][Vh][4][3]         parity 0 first
 * lattice_size is {Lx, Ly, Lz, Lt}; Lx must be even.  One GPU holds the whole
 * lattice, so grid_size must be {1,1,1,1}.
 *
 * The operators use the reference normalisation: kappa = 1, c_sw * kappa = 1,
 * periodic boundaries.  The extended API (qcu_ex.h) exposes kappa, c_sw, the t
 * boundary, precision, multiple right-hand sides and solve statistics.
 *
 * Errors: these functions return void as in the reference.  A failure prints a
 * message to stderr, records it for qcuGetLastError()/qcuGetLastErrorString()
 * (qcu_ex.h) and, unless the environment sets QCU_EXIT_ON_ERROR=0, ends the
 * process with exit(-1) as the reference's checkCudaErrors does
 * (include/qcu_macro.cuh:34-41).
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

/* include/qcu.h:7-9 */
typedef struct QcuParam_s {
  int lattice_size[4];
} QcuParam;

/* include/qcu.h:11-13 */
typedef struct QcuGrid_t {
  int grid_size[4];
} QcuGrid_t;

/* Replaces initGridSize, include/qcu.h:14 (src/qcu_communicator.cu:1630-1650).
 * Records the lattice geometry and the gauge pointer; no MPI.  grid_size other than
 * {1,1,1,1} is an error (this build keeps the whole lattice on one GPU). */
void initGridSize(QcuGrid_t* grid, QcuParam* p_param, void* gauge, void* fermion_in, void* fermion_out);

/* Replaces dslashQcu, include/qcu.h:15 (src/dslash.cu:62-77).
 * fermion_out (half spinor of parity `parity`) = D fermion_in (half spinor of parity
 * 1 - parity): the hopping term with s = +1, kappa = 1, periodic boundaries. */
void dslashQcu(void* fermion_out, void* fermion_in, void* gauge, QcuParam* param, int parity);

/* Replaces fullDslashQcu, include/qcu.h:17 (src/qcu_clover_dslash.cu:948-987).
 * Full even-first spinors: fermion_out = (A - D_s) fermion_in with A = 1 + T the
 * clover term built from `gauge`, s = -1 if dagger_flag else +1. */
void fullDslashQcu(void* fermion_out, void* fermion_in, void* gauge, QcuParam* param, int dagger_flag);

/* Replaces cg_inverter, include/qcu.h:18 (src/cg_inverter.cu:260-369).
 * Full spinors: x_vector <- (A - D)^-1 b_vector; b_vector is not modified.  Solved
 * by even-odd preconditioned CG on the normal equations to a true relative residual
 * ||b - M x|| / ||b|| of 1e-12 within 20 Vh iterations (qcuSetDefaultSolve in
 * qcu_ex.h changes both).  As in the reference, a solve that does not converge is
 * not fatal: x holds the last iterate, "cg_failed!!!" is printed and
 * qcuGetLastError() returns QCU_ERROR_NOT_CONVERGED. */
void cg_inverter(void* b_vector, void* x_vector, void* gauge, QcuParam* param);

/* Replaces loadQcuGauge, include/qcu.h:20 (src/dslash.cu:27-35).
 * Packs `gauge` into the library's resident layout; later calls that pass the same
 * pointer use the packed copy until loadQcuGauge is called on it again. */
void loadQcuGauge(void* gauge, QcuParam* param);

#ifdef __cplusplus
}
#endif
