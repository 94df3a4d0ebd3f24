/*
 * qcu_oracle.c -- CPU FP64 restatement of the qcu even-odd Wilson / Wilson-clover
 * arithmetic.  TEST INFRASTRUCTURE ONLY: the checker for libqcu.so, never part of it.
 *
 * Reference files followed (paths relative to cg_inverter/ of the reference):
 *   complex product / divide  include/qcu_complex.cuh:44-67
 *   site moves                include/qcu_point.cuh:20-63
 *   reconstruct-12            src/qcu_wilson_dslash_neo.cu:73-78
 *   hopping term              src/qcu_wilson_dslash_neo.cu:941-1206 (s = +1),
 *                             src/qcu_clover_dslash.cu:570-772 (s = +-1)
 *   clover leaves / T         src/qcu_clover_dslash.cu:173-267, 329-458
 *   A = 1 + T, inverse        src/qcu_clover_dslash.cu:42-100, 509-556
 *   clover block mat-vec      src/qcu_clover_dslash.cu:461-506
 *   operator compositions     src/qcu_clover_dslash.cu:902-1049
 *   CG                        src/cg_inverter.cu:125-369
 *   inner product convention  src/qcu_complex_computation.cu:44-72 (sum conj(x) y)
 * Extensions that the reference does not have (kappa, c_sw, t boundary sign, the
 * even-odd Schur system, Cholesky, the seeded generators) are written from the
 * mathematics and pinned by the tests listed in qcu_oracle.h.
 *
 * Build with -ffp-contract=off so every product and sum is rounded separately.
 */
#include "qcu_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cpx;

static inline cpx C(double r, double i) { cpx z; z.re = r; z.im = i; return z; }
static inline cpx cadd(cpx a, cpx b) { return C(a.re + b.re, a.im + b.im); }
static inline cpx csub(cpx a, cpx b) { return C(a.re - b.re, a.im - b.im); }
/* include/qcu_complex.cuh:44-47 */
static inline cpx cmul(cpx a, cpx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cpx cscale(cpx a, double s) { return C(a.re * s, a.im * s); }
static inline cpx cconj(cpx a) { return C(a.re, -a.im); }
/* multiply by a phase in {1, -1, i, -i}: exact */
static inline cpx cphase(cpx a, int ph) {
  switch (ph & 3) {
    case 0: return a;
    case 1: return C(-a.im, a.re);   /* i,  include/qcu_complex.cuh:97-100 */
    case 2: return C(-a.re, -a.im);  /* -1 */
    default: return C(a.im, -a.re);  /* -i, include/qcu_complex.cuh:101-104 */
  }
}
/* a / b = a conj(b) / |b|^2, include/qcu_complex.cuh:65-67 */
static inline cpx cdiv(cpx a, cpx b) {
  double d = b.re * b.re + b.im * b.im;
  cpx n = cmul(a, cconj(b));
  return C(n.re / d, n.im / d);
}

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
static int nthreads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* geometry                                                                   */
/* ------------------------------------------------------------------------- */
void orc_coords(const int L[4], int parity, long idx, int c[4]) {
  int lxh = L[0] / 2;
  long r = idx;
  int xh = (int)(r % lxh); r /= lxh;
  int y = (int)(r % L[1]); r /= L[1];
  int z = (int)(r % L[2]); r /= L[2];
  int t = (int)r;
  c[0] = 2 * xh + ((parity + y + z + t) & 1);
  c[1] = y; c[2] = z; c[3] = t;
}

long orc_index(const int L[4], const int cin[4], int* parity) {
  int c[4];
  for (int m = 0; m < 4; ++m) c[m] = ((cin[m] % L[m]) + L[m]) % L[m];
  if (parity) *parity = (c[0] + c[1] + c[2] + c[3]) & 1;
  return (((long)c[3] * L[2] + c[2]) * L[1] + c[1]) * (L[0] / 2) + c[0] / 2;
}

/* include/qcu_point.cuh:20-63: x moves only when (y+z+t)&1 allows it */
long orc_neighbor(const int L[4], int parity, long idx, int mu, int fwd, int* wrapped) {
  int lxh = L[0] / 2;
  long r = idx;
  int xh = (int)(r % lxh); r /= lxh;
  int y = (int)(r % L[1]); r /= L[1];
  int z = (int)(r % L[2]); r /= L[2];
  int t = (int)r;
  int eo = (y + z + t) & 1;
  int w = 0;
  switch (mu) {
    case 0:
      if (fwd) {
        if (eo != parity) { if (xh == lxh - 1) { xh = 0; w = 1; } else ++xh; }
      } else {
        if (eo == parity) { if (xh == 0) { xh = lxh - 1; w = 1; } else --xh; }
      }
      break;
    case 1:
      if (fwd) { if (y == L[1] - 1) { y = 0; w = 1; } else ++y; }
      else { if (y == 0) { y = L[1] - 1; w = 1; } else --y; }
      break;
    case 2:
      if (fwd) { if (z == L[2] - 1) { z = 0; w = 1; } else ++z; }
      else { if (z == 0) { z = L[2] - 1; w = 1; } else --z; }
      break;
    default:
      if (fwd) { if (t == L[3] - 1) { t = 0; w = 1; } else ++t; }
      else { if (t == 0) { t = L[3] - 1; w = 1; } else --t; }
      break;
  }
  if (wrapped) *wrapped = w;
  return (((long)t * L[2] + z) * L[1] + y) * lxh + xh;
}

/* ------------------------------------------------------------------------- */
/* SU(3) helpers                                                              */
/* ------------------------------------------------------------------------- */
static void recon12(cpx* u) {
  /* src/qcu_wilson_dslash_neo.cu:73-78 */
  u[6] = cconj(csub(cmul(u[1], u[5]), cmul(u[2], u[4])));
  u[7] = cconj(csub(cmul(u[2], u[3]), cmul(u[0], u[5])));
  u[8] = cconj(csub(cmul(u[0], u[4]), cmul(u[1], u[3])));
}
void orc_reconstruct12(double* u) { recon12((cpx*)u); }

static inline long vh_of(const int L[4]) { return (long)L[0] * L[1] * L[2] * L[3] / 2; }

/* ------------------------------------------------------------------------- */
/* layouts                                                                    */
/* ------------------------------------------------------------------------- */
/* src/qcu_shift_storage.cu:74-116,166-180: one half spinor, AoS [Vh][24] doubles
 * <-> coalesced [24][Vh]; dir 0 = to coalesced */
void orc_shift_vector(double* dst, const double* src, int dir, long vh) {
  for (long s = 0; s < vh; ++s)
    for (int i = 0; i < 24; ++i) {
      if (dir == 0) dst[i * vh + s] = src[s * 24 + i];
      else dst[s * 24 + i] = src[i * vh + s];
    }
}

/* src/qcu_shift_storage.cu:123-161,183-194: per (mu, parity) field, links [Vh][18]
 * doubles -> rows 0-1 as [12][Vh] (dir 0).  dir 1 (left empty by the reference) goes
 * back and rebuilds row 2 as recon12 does. */
void orc_shift_gauge(double* dst, const double* src, int dir, long vh) {
  for (int f = 0; f < 8; ++f)
    for (long s = 0; s < vh; ++s) {
      if (dir == 0) {
        const double* a = src + ((long)f * vh + s) * 18;
        for (int j = 0; j < 12; ++j) dst[((long)f * 12 + j) * vh + s] = a[j];
      } else {
        cpx u[9];
        double* d = (double*)u;
        for (int j = 0; j < 12; ++j) d[j] = src[((long)f * 12 + j) * vh + s];
        recon12(u);
        memcpy(dst + ((long)f * vh + s) * 18, u, sizeof u);
      }
    }
}

/* lexicographic site ((t*Lz+z)*Ly+y)*Lx+x <-> even-odd (parity, half-site index) of
 * orc_coords; nf fields of nc complex per site.  to_eo: lex -> even-odd. */
void orc_lex_eo(double* dst, const double* src, const int L[4], int nc, int nf, int to_eo) {
  long vh = vh_of(L);
  const cpx* a = (const cpx*)src;
  cpx* b = (cpx*)dst;
  for (int f = 0; f < nf; ++f)
    for (int p = 0; p < 2; ++p)
      for (long i = 0; i < vh; ++i) {
        int c[4];
        orc_coords(L, p, i, c);
        long l = (((long)c[3] * L[2] + c[2]) * L[1] + c[1]) * L[0] + c[0];
        long e = (long)f * 2 * vh + p * vh + i;
        long x = (long)f * 2 * vh + l;
        for (int k = 0; k < nc; ++k) {
          if (to_eo) b[e * nc + k] = a[x * nc + k];
          else b[x * nc + k] = a[e * nc + k];
        }
      }
}

/* link U_mu at (parity, idx); include/qcu_point.cuh:65-67 */
static void load_link(cpx* u, const double* gauge, const int L[4], int mu, int parity, long idx, int recon) {
  const cpx* g = (const cpx*)gauge + ((((long)mu * 2 + parity) * vh_of(L)) + idx) * 9;
  if (recon == 12) {
    for (int i = 0; i < 6; ++i) u[i] = g[i];
    recon12(u);
  } else {
    for (int i = 0; i < 9; ++i) u[i] = g[i];
  }
}

/* ------------------------------------------------------------------------- */
/* hopping term                                                               */
/* ------------------------------------------------------------------------- */
/* gamma_mu (DeGrand-Rossi as read off the projections in
 * src/qcu_wilson_dslash_neo.cu:409-416,468-475,524-531,581-588): each row has one
 * nonzero entry, at column gcol[mu][row] with phase gph[mu][row]
 * (0 = 1, 1 = i, 2 = -1, 3 = -i). */
static const int gcol[4][4] = {{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}, {2, 3, 0, 1}};
static const int gph[4][4] = {{1, 1, 3, 3}, {2, 0, 0, 2}, {1, 3, 3, 1}, {0, 0, 0, 0}};

/* one hop direction: acc += (U or U^dag) (1 + sigma g_mu) psi, sigma = -s forward,
 * +s backward (src/qcu_wilson_dslash_neo.cu:974-1193 per direction).  The spin
 * projection keeps the two upper components; the lower two are +-(i) multiples
 * of them. */
static void hop_dir(cpx* acc, const cpx* u, int dag, const cpx* psi, int mu, int sigma, double bsign) {
  int sph = sigma > 0 ? 0 : 2;
  cpx h[2][3];
  for (int c = 0; c < 3; ++c) {
    h[0][c] = cadd(psi[0 * 3 + c], cphase(psi[gcol[mu][0] * 3 + c], gph[mu][0] + sph));
    h[1][c] = cadd(psi[1 * 3 + c], cphase(psi[gcol[mu][1] * 3 + c], gph[mu][1] + sph));
    if (bsign != 1.0) { h[0][c] = cscale(h[0][c], bsign); h[1][c] = cscale(h[1][c], bsign); }
  }
  for (int i = 0; i < 3; ++i) {
    cpx t0 = C(0, 0), t1 = C(0, 0);
    for (int j = 0; j < 3; ++j) {
      cpx uij = dag ? cconj(u[j * 3 + i]) : u[i * 3 + j];
      t0 = cadd(t0, cmul(h[0][j], uij));
      t1 = cadd(t1, cmul(h[1][j], uij));
    }
    cpx th[2] = {t0, t1};
    acc[0 * 3 + i] = cadd(acc[0 * 3 + i], t0);
    acc[1 * 3 + i] = cadd(acc[1 * 3 + i], t1);
    /* rows 2,3 of (1 + sigma g) psi equal sigma g[row][col] h[col] */
    acc[2 * 3 + i] = cadd(acc[2 * 3 + i], cphase(th[gcol[mu][2]], gph[mu][2] + sph));
    acc[3 * 3 + i] = cadd(acc[3 * 3 + i], cphase(th[gcol[mu][3]], gph[mu][3] + sph));
  }
}

static void hop_site(cpx* acc, const cpx* in, const double* gauge, const int L[4], int parity,
                     long idx, int s, double tbc, int recon) {
  int c[4];
  orc_coords(L, parity, idx, c);
  cpx u[9];
  for (int k = 0; k < 12; ++k) acc[k] = C(0, 0);
  for (int mu = 0; mu < 4; ++mu) {
    int w;
    /* forward: U_mu(x) (1 - s g_mu) psi(x+mu) */
    long nf = orc_neighbor(L, parity, idx, mu, 1, &w);
    double bs = (mu == 3 && w) ? tbc : 1.0;
    load_link(u, gauge, L, mu, parity, idx, recon);
    hop_dir(acc, u, 0, in + nf * 12, mu, -s, bs);
    /* backward: U_mu(x-mu)^dag (1 + s g_mu) psi(x-mu) */
    long nb = orc_neighbor(L, parity, idx, mu, 0, &w);
    bs = (mu == 3 && w) ? tbc : 1.0;
    load_link(u, gauge, L, mu, 1 - parity, nb, recon);
    hop_dir(acc, u, 1, in + nb * 12, mu, s, bs);
  }
}

void orc_hop(double* out, const double* in, const double* gauge, const int L[4], int parity,
             int s, double tbc, int recon) {
  long vh = vh_of(L);
  const cpx* cin = (const cpx*)in;
  cpx* cout = (cpx*)out;
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long i = 0; i < vh; ++i) {
    cpx acc[12];
    hop_site(acc, cin, gauge, L, parity, i, s, tbc, recon);
    for (int k = 0; k < 12; ++k) cout[i * 12 + k] = acc[k];
  }
}

/* ------------------------------------------------------------------------- */
/* clover                                                                     */
/* ------------------------------------------------------------------------- */
static void su3_mul(cpx* r, const cpx* a, const cpx* b) { /* src/qcu_clover_dslash.cu:173-185 */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cpx t = C(0, 0);
      for (int k = 0; k < 3; ++k) t = cadd(t, cmul(a[i * 3 + k], b[k * 3 + j]));
      r[i * 3 + j] = t;
    }
}
static void su3_dag(cpx* r, const cpx* a) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i * 3 + j] = cconj(a[j * 3 + i]);
}
/* link U_mu at full coordinates c + d (periodic) */
static void link_at(cpx* u, const double* gauge, const int L[4], const int c[4], const int d[4], int mu,
                    int recon) {
  int cc[4] = {c[0] + d[0], c[1] + d[1], c[2] + d[2], c[3] + d[3]};
  int p;
  long idx = orc_index(L, cc, &p);
  load_link(u, gauge, L, mu, p, idx, recon);
}

/* Gamma_munu = g_mu g_nu restricted to its nonzero entries, as tabulated in
 * src/qcu_clover_dslash.cu:210-267: for plane k, entry e: (row spin, col spin, phase) */
static const int cl_row[6][4] = {{0, 1, 2, 3}, {0, 1, 2, 3}, {0, 1, 2, 3}, {0, 1, 2, 3}, {0, 1, 2, 3}, {0, 1, 2, 3}};
static const int cl_col[6][4] = {{0, 1, 2, 3}, {1, 0, 3, 2}, {1, 0, 3, 2}, {1, 0, 3, 2}, {1, 0, 3, 2}, {0, 1, 2, 3}};
static const int cl_ph[6][4] = {{3, 1, 3, 1}, {2, 0, 2, 0}, {1, 1, 3, 3}, {3, 3, 3, 3}, {2, 0, 0, 2}, {1, 3, 3, 1}};
static const int plane_mu[6] = {0, 0, 0, 1, 1, 2};
static const int plane_nu[6] = {1, 2, 3, 2, 3, 3};

static void clover_T_site(cpx* T, const double* gauge, const int L[4], int parity, long idx, int recon) {
  int c[4];
  orc_coords(L, parity, idx, c);
  for (int k = 0; k < 72; ++k) T[k] = C(0, 0);
  for (int pl = 0; pl < 6; ++pl) {
    int mu = plane_mu[pl], nu = plane_nu[pl];
    int z[4] = {0, 0, 0, 0}, pm[4] = {0, 0, 0, 0}, pn[4] = {0, 0, 0, 0}, mm[4] = {0, 0, 0, 0},
        mn[4] = {0, 0, 0, 0}, mmpn[4] = {0, 0, 0, 0}, mmmn[4] = {0, 0, 0, 0}, mnpm[4] = {0, 0, 0, 0};
    pm[mu] = 1; pn[nu] = 1; mm[mu] = -1; mn[nu] = -1;
    mmpn[mu] = -1; mmpn[nu] = 1;
    mmmn[mu] = -1; mmmn[nu] = -1;
    mnpm[nu] = -1; mnpm[mu] = 1;
    cpx a[9], b[9], d[9], t1[9], t2[9], sum[9];
    for (int k = 0; k < 9; ++k) sum[k] = C(0, 0);
    /* F_1 = U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag   (:329-350) */
    link_at(a, gauge, L, c, z, mu, recon);
    link_at(b, gauge, L, c, pm, nu, recon);
    su3_mul(t1, a, b);
    link_at(a, gauge, L, c, pn, mu, recon); su3_dag(d, a);
    su3_mul(t2, t1, d);
    link_at(a, gauge, L, c, z, nu, recon); su3_dag(d, a);
    su3_mul(t1, t2, d);
    for (int k = 0; k < 9; ++k) sum[k] = cadd(sum[k], t1[k]);
    /* F_2 = U_nu(x) U_mu(x-mu+nu)^dag U_nu(x-mu)^dag U_mu(x-mu)   (:351-372) */
    link_at(a, gauge, L, c, z, nu, recon);
    link_at(b, gauge, L, c, mmpn, mu, recon); su3_dag(d, b);
    su3_mul(t1, a, d);
    link_at(b, gauge, L, c, mm, nu, recon); su3_dag(d, b);
    su3_mul(t2, t1, d);
    link_at(b, gauge, L, c, mm, mu, recon);
    su3_mul(t1, t2, b);
    for (int k = 0; k < 9; ++k) sum[k] = cadd(sum[k], t1[k]);
    /* F_3 = U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu)   (:373-392) */
    link_at(b, gauge, L, c, mm, mu, recon); su3_dag(a, b);
    link_at(b, gauge, L, c, mmmn, nu, recon); su3_dag(d, b);
    su3_mul(t1, a, d);
    link_at(b, gauge, L, c, mmmn, mu, recon);
    su3_mul(t2, t1, b);
    link_at(b, gauge, L, c, mn, nu, recon);
    su3_mul(t1, t2, b);
    for (int k = 0; k < 9; ++k) sum[k] = cadd(sum[k], t1[k]);
    /* F_4 = U_nu(x-nu)^dag U_mu(x-nu) U_nu(x-nu+mu) U_mu(x)^dag   (:393-413) */
    link_at(b, gauge, L, c, mn, nu, recon); su3_dag(a, b);
    link_at(b, gauge, L, c, mn, mu, recon);
    su3_mul(t1, a, b);
    link_at(b, gauge, L, c, mnpm, nu, recon);
    su3_mul(t2, t1, b);
    link_at(b, gauge, L, c, z, mu, recon); su3_dag(d, b);
    su3_mul(t1, t2, d);
    for (int k = 0; k < 9; ++k) sum[k] = cadd(sum[k], t1[k]);
    /* G = Q - Q^dag (:440-444) */
    cpx g[9];
    su3_dag(d, sum);
    for (int k = 0; k < 9; ++k) g[k] = csub(sum[k], d[k]);
    /* T += (-1/16) * phase * G * 2 into the chiral blocks (:210-267, :446) */
    cpx coef = C(-1.0 / 16.0, 0.0);
    for (int e = 0; e < 4; ++e) {
      int rs = cl_row[pl][e], cs = cl_col[pl][e];
      int blk = rs / 2;
      cpx ph = cphase(C(1, 0), cl_ph[pl][e]);
      cpx cph = cmul(coef, ph);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          int row = (rs % 2) * 3 + i, col = (cs % 2) * 3 + j;
          cpx v = cscale(cmul(cph, g[i * 3 + j]), 2.0);
          T[blk * 36 + row * 6 + col] = cadd(T[blk * 36 + row * 6 + col], v);
        }
    }
  }
}

void orc_clover_T(double* T, const double* gauge, const int L[4], int recon) {
  long vh = vh_of(L);
  cpx* ct = (cpx*)T;
  for (int p = 0; p < 2; ++p) {
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long i = 0; i < vh; ++i) clover_T_site(ct + ((long)p * vh + i) * 72, gauge, L, p, i, recon);
  }
}

void orc_clover_A(double* A, const double* T, long nsites, double ck) {
  const cpx* ct = (const cpx*)T;
  cpx* ca = (cpx*)A;
  for (long s = 0; s < nsites; ++s) {
    for (int k = 0; k < 72; ++k) ca[s * 72 + k] = cscale(ct[s * 72 + k], ck);
    for (int b = 0; b < 2; ++b)
      for (int i = 0; i < 6; ++i) {
        cpx* d = &ca[s * 72 + b * 36 + i * 6 + i];
        *d = cadd(*d, C(1, 0));   /* :537-540 */
      }
  }
}

/* src/qcu_clover_dslash.cu:42-100: augmented [M | 1], no pivoting unless a
 * diagonal entry is exactly zero, row divide then eliminate every other row */
void orc_invert6_gj(const double* m, double* inv) {
  const cpx* cm = (const cpx*)m;
  cpx aug[6][12];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 12; ++j) aug[i][j] = j < 6 ? cm[i * 6 + j] : C(i == j - 6 ? 1 : 0, 0);
  for (int i = 0; i < 6; ++i) {
    if (aug[i][i].re == 0 && aug[i][i].im == 0) {
      for (int j = i + 1; j < 6; ++j)
        if (aug[j][i].re != 0 || aug[j][i].im != 0) {
          for (int k = 0; k < 12; ++k) { cpx t = aug[i][k]; aug[i][k] = aug[j][k]; aug[j][k] = t; }
          break;
        }
    }
    if (!(aug[i][i].re == 1 && aug[i][i].im == 0)) {
      cpx num = aug[i][i];
      double d = num.re * num.re + num.im * num.im;
      for (int k = 0; k < 12; ++k) {  /* operator/=, include/qcu_complex.cuh:68-74 */
        double nr = (aug[i][k].re * num.re + aug[i][k].im * num.im) / d;
        double ni = (num.re * aug[i][k].im - aug[i][k].re * num.im) / d;
        aug[i][k] = C(nr, ni);
      }
    }
    for (int j = 0; j < 6; ++j) {
      if (j == i) continue;
      cpx num = aug[j][i];
      for (int k = 0; k < 12; ++k) aug[j][k] = csub(aug[j][k], cmul(num, aug[i][k]));
    }
  }
  cpx* ci = (cpx*)inv;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) ci[i * 6 + j] = aug[i][j + 6];
}

/* Hermitian positive-definite inverse: M = L L^dag, M^-1 = L^-dag L^-1.  Reads the
 * lower triangle of M only. */
int orc_invert6_chol(const double* m, double* inv) {
  const cpx* cm = (const cpx*)m;
  cpx l[6][6];
  double dinv[6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) l[i][j] = C(0, 0);
  for (int j = 0; j < 6; ++j) {
    double d = cm[j * 6 + j].re;
    for (int k = 0; k < j; ++k) d -= l[j][k].re * l[j][k].re + l[j][k].im * l[j][k].im;
    if (!(d > 0.0)) return -1;
    double lj = sqrt(d);
    l[j][j] = C(lj, 0);
    dinv[j] = 1.0 / lj;
    for (int i = j + 1; i < 6; ++i) {
      cpx s = cm[i * 6 + j];
      for (int k = 0; k < j; ++k) s = csub(s, cmul(l[i][k], cconj(l[j][k])));
      l[i][j] = cscale(s, dinv[j]);
    }
  }
  /* W = L^-1 (lower triangular) */
  cpx w[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) w[i][j] = C(0, 0);
  for (int j = 0; j < 6; ++j) {
    w[j][j] = C(dinv[j], 0);
    for (int i = j + 1; i < 6; ++i) {
      cpx s = C(0, 0);
      for (int k = j; k < i; ++k) s = cadd(s, cmul(l[i][k], w[k][j]));
      w[i][j] = cscale(s, -dinv[i]);
    }
  }
  /* M^-1_ij = sum_{k >= max(i,j)} conj(W_ki) W_kj */
  cpx* ci = (cpx*)inv;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      cpx s = C(0, 0);
      for (int k = (i > j ? i : j); k < 6; ++k) s = cadd(s, cmul(cconj(w[k][i]), w[k][j]));
      ci[i * 6 + j] = s;
    }
  return 0;
}

int orc_clover_invert(double* Ainv, const double* A, long nsites, int chol) {
  int bad = 0;
#pragma omp parallel for schedule(static) num_threads(nthreads()) reduction(+ : bad)
  for (long s = 0; s < nsites; ++s)
    for (int b = 0; b < 2; ++b) {
      const double* m = A + (s * 72 + b * 36) * 2;
      double* r = Ainv + (s * 72 + b * 36) * 2;
      if (chol) {
        if (orc_invert6_chol(m, r) != 0) { orc_invert6_gj(m, r); bad++; }
      } else {
        orc_invert6_gj(m, r);
      }
    }
  return bad;
}

void orc_clover_apply(double* out, const double* in, const double* M, long vh) {
  const cpx* ci = (const cpx*)in;
  const cpx* cmx = (const cpx*)M;
  cpx* co = (cpx*)out;
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long s = 0; s < vh; ++s) {
    cpx src[12], dst[12];
    for (int k = 0; k < 12; ++k) { src[k] = ci[s * 12 + k]; dst[k] = C(0, 0); }
    for (int b = 0; b < 2; ++b)
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
          dst[b * 6 + i] = cadd(dst[b * 6 + i], cmul(cmx[s * 72 + b * 36 + i * 6 + j], src[b * 6 + j]));
    for (int k = 0; k < 12; ++k) co[s * 12 + k] = dst[k];
  }
}

/* ------------------------------------------------------------------------- */
/* operators                                                                  */
/* ------------------------------------------------------------------------- */
static void axpy_c(double* y, double a, const double* x, long n) { for (long i = 0; i < n; ++i) y[i] += a * x[i]; }

void orc_full_op(double* out, const double* in, const double* gauge, const double* A, const int L[4],
                 double kappa, int s, double tbc, int recon) {
  long vh = vh_of(L);
  long hn = vh * 24;
  double* hop = (double*)malloc(sizeof(double) * hn);
  for (int p = 0; p < 2; ++p) {
    orc_hop(hop, in + (1 - p) * hn, gauge, L, p, s, tbc, recon);
    if (A) orc_clover_apply(out + p * hn, in + p * hn, A + (long)p * vh * 144, vh);
    else memcpy(out + p * hn, in + p * hn, sizeof(double) * hn);
    axpy_c(out + p * hn, -kappa, hop, hn);
  }
  free(hop);
}

void orc_schur_op(double* out, const double* in, const double* gauge, const double* Aee,
                  const double* Ainv_oo, const int L[4], double kappa, int s, double tbc, int recon) {
  long vh = vh_of(L);
  long hn = vh * 24;
  double* t1 = (double*)malloc(sizeof(double) * hn);
  double* t2 = (double*)malloc(sizeof(double) * hn);
  orc_hop(t1, in, gauge, L, 1, s, tbc, recon);
  if (Ainv_oo) { orc_clover_apply(t2, t1, Ainv_oo, vh); memcpy(t1, t2, sizeof(double) * hn); }
  orc_hop(t2, t1, gauge, L, 0, s, tbc, recon);
  if (Aee) orc_clover_apply(out, in, Aee, vh);
  else memcpy(out, in, sizeof(double) * hn);
  axpy_c(out, -kappa * kappa, t2, hn);
  free(t1);
  free(t2);
}

double orc_norm2(const double* v, long n_complex) {
  double s = 0;
  for (long i = 0; i < 2 * n_complex; ++i) s += v[i] * v[i];
  return s;
}
void orc_cdot(const double* x, const double* y, long n_complex, double* re, double* im) {
  cpx acc = C(0, 0);
  const cpx* cx = (const cpx*)x;
  const cpx* cy = (const cpx*)y;
  for (long i = 0; i < n_complex; ++i) acc = cadd(acc, cmul(cconj(cx[i]), cy[i]));
  *re = acc.re;
  *im = acc.im;
}

/* Even-odd CG on the normal equations in CGLS form (tracks s = b^ - Mhat x so the
 * stopping test is on the residual of the system itself):
 *   s = b^, r = Mhat^dag s, p = r
 *   t = Mhat p; a = |r|^2/|t|^2; x += a p; s -= a t; r = Mhat^dag s; b = |r_new|^2/|r|^2; p = r + b p
 * Stops when |s| <= tol |b| (|b| of the full-lattice source). */
int orc_cg_eo(double* x, const double* b, const double* gauge, const double* A, const double* Ainv,
              const int L[4], double kappa, double tbc, int recon, double tol, int max_iter,
              OrcSolveStats* st) {
  long vh = vh_of(L);
  long hn = vh * 24;
  const double* Aee = A ? A : NULL;
  const double* Aoo_inv = Ainv ? Ainv + vh * 144 : NULL;
  double* bh = (double*)malloc(sizeof(double) * hn);
  double* s = (double*)malloc(sizeof(double) * hn);
  double* r = (double*)malloc(sizeof(double) * hn);
  double* p = (double*)malloc(sizeof(double) * hn);
  double* t = (double*)malloc(sizeof(double) * hn);
  double* tmp = (double*)malloc(sizeof(double) * hn);
  double* xe = x;
  double* xo = x + hn;
  const double* be = b;
  const double* bo = b + hn;

  /* b^_e = b_e + kappa D_eo A_oo^-1 b_o */
  if (Aoo_inv) orc_clover_apply(tmp, bo, Aoo_inv, vh);
  else memcpy(tmp, bo, sizeof(double) * hn);
  orc_hop(bh, tmp, gauge, L, 0, +1, tbc, recon);
  for (long i = 0; i < hn; ++i) bh[i] = be[i] + kappa * bh[i];

  double bnorm = sqrt(orc_norm2(b, 2 * vh * 12));
  double target = tol * bnorm;
  memset(xe, 0, sizeof(double) * hn);
  memcpy(s, bh, sizeof(double) * hn);
  orc_schur_op(r, s, gauge, Aee, Aoo_inv, L, kappa, -1, tbc, recon);
  memcpy(p, r, sizeof(double) * hn);
  double rr = orc_norm2(r, vh * 12);
  double ss = orc_norm2(s, vh * 12);
  int it = 0;
  int conv = sqrt(ss) <= target;
  while (!conv && it < max_iter) {
    orc_schur_op(t, p, gauge, Aee, Aoo_inv, L, kappa, +1, tbc, recon);
    double tt = orc_norm2(t, vh * 12);
    double a = rr / tt;
    for (long i = 0; i < hn; ++i) { xe[i] += a * p[i]; s[i] -= a * t[i]; }
    ss = orc_norm2(s, vh * 12);
    orc_schur_op(r, s, gauge, Aee, Aoo_inv, L, kappa, -1, tbc, recon);
    double rr_new = orc_norm2(r, vh * 12);
    double beta = rr_new / rr;
    rr = rr_new;
    for (long i = 0; i < hn; ++i) p[i] = r[i] + beta * p[i];
    ++it;
    conv = sqrt(ss) <= target;
  }
  /* x_o = A_oo^-1 (b_o + kappa D_oe x_e) */
  orc_hop(tmp, xe, gauge, L, 1, +1, tbc, recon);
  for (long i = 0; i < hn; ++i) tmp[i] = bo[i] + kappa * tmp[i];
  if (Aoo_inv) orc_clover_apply(xo, tmp, Aoo_inv, vh);
  else memcpy(xo, tmp, sizeof(double) * hn);

  if (st) {
    double* full = (double*)malloc(sizeof(double) * 2 * hn);
    orc_full_op(full, x, gauge, A, L, kappa, +1, tbc, recon);
    for (long i = 0; i < 2 * hn; ++i) full[i] = b[i] - full[i];
    st->iterations = it;
    st->true_residual = sqrt(orc_norm2(full, 2 * vh * 12)) / bnorm;
    st->normal_residual = sqrt(ss) / bnorm;
    st->converged = conv;
    free(full);
  }
  free(bh); free(s); free(r); free(p); free(t); free(tmp);
  return conv ? 0 : 1;
}

/* M psi = psi - A^-1 D^s psi on the full lattice (newFullCloverDslashOneRound,
 * src/qcu_clover_dslash.cu:1013-1042; kappa = 1, periodic, reconstructed links) */
static void ref_m(double* out, const double* in, const double* gauge, const double* Ainv, const int L[4],
                  int s) {
  long vh = vh_of(L);
  long hn = vh * 24;
  double* tmp = (double*)malloc(sizeof(double) * hn);
  for (int p = 0; p < 2; ++p) {
    orc_hop(tmp, in + (1 - p) * hn, gauge, L, p, s, 1.0, 12);
    if (Ainv) orc_clover_apply(out + p * hn, tmp, Ainv + (long)p * vh * 144, vh);
    else memcpy(out + p * hn, tmp, sizeof(double) * hn);
    for (long i = 0; i < hn; ++i) out[p * hn + i] = in[p * hn + i] - out[p * hn + i];
  }
  free(tmp);
}

static void cdot_c(const double* x, const double* y, long n, cpx* r) { orc_cdot(x, y, n, &r->re, &r->im); }

/* src/cg_inverter.cu:260-369 and cg() :181-256, complex alpha/beta divided as
 * in include/qcu_complex.cuh:65-67 */
int orc_cg_ref(double* x, const double* b, const double* gauge, const double* Ainv, const int L[4],
               double stop_rr, int max_iter, int* iters) {
  long n = vh_of(L) * 2 * 12; /* complex entries of a full spinor */
  long nd = 2 * n;
  double* nb = (double*)malloc(sizeof(double) * nd);
  double* r = (double*)malloc(sizeof(double) * nd);
  double* p = (double*)malloc(sizeof(double) * nd);
  double* q = (double*)malloc(sizeof(double) * nd);
  double* tmp = (double*)malloc(sizeof(double) * nd);
  double* r2 = (double*)malloc(sizeof(double) * nd);
  long vh = vh_of(L);
  memset(x, 0, sizeof(double) * nd);
  /* new_b = M' A^-1 b */
  if (Ainv) {
    orc_clover_apply(tmp, b, Ainv, vh);
    orc_clover_apply(tmp + vh * 24, b + vh * 24, Ainv + vh * 144, vh);
  } else {
    memcpy(tmp, b, sizeof(double) * nd);
  }
  ref_m(nb, tmp, gauge, Ainv, L, -1);
  memcpy(r, nb, sizeof(double) * nd); /* x = 0 so r = new_b */
  memcpy(p, r, sizeof(double) * nd);
  cpx rr;
  cdot_c(r, r, n, &rr);
  int it = 0, done = sqrt(rr.re * rr.re + rr.im * rr.im) < stop_rr;
  while (!done && it < max_iter) {
    cpx num, den;
    cdot_c(r, r, n, &num);
    ref_m(tmp, p, gauge, Ainv, L, +1);
    ref_m(q, tmp, gauge, Ainv, L, -1);
    cdot_c(p, q, n, &den);
    cpx alpha = cdiv(num, den);
    for (long i = 0; i < n; ++i) {
      cpx* xi = (cpx*)x + i;
      *xi = cadd(*xi, cmul(((cpx*)p)[i], alpha));
    }
    cpx malpha = cmul(alpha, C(-1, 0));
    memcpy(r2, r, sizeof(double) * nd);
    for (long i = 0; i < n; ++i) {
      cpx* ri = (cpx*)r2 + i;
      *ri = cadd(*ri, cmul(((cpx*)q)[i], malpha));
    }
    ++it;
    /* convergence is tested on the previous r (src/cg_inverter.cu:225) */
    if (sqrt(num.re * num.re + num.im * num.im) < stop_rr) { done = 1; break; }
    cdot_c(r2, r2, n, &den);
    cpx beta = cdiv(den, num);
    for (long i = 0; i < n; ++i) {
      cpx* pi = (cpx*)p + i;
      *pi = cadd(cmul(beta, *pi), ((cpx*)r2)[i]);
    }
    memcpy(r, r2, sizeof(double) * nd);
  }
  if (iters) *iters = it;
  free(nb); free(r); free(p); free(q); free(tmp); free(r2);
  return done ? 0 : 1;
}

/* ------------------------------------------------------------------------- */
/* seeded inputs                                                              */
/* ------------------------------------------------------------------------- */
static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
uint64_t orc_rng_u64(uint64_t seed, uint64_t ctr) { return mix64(mix64(seed) ^ ctr); }
static inline double rng_uniform(uint64_t seed, uint64_t ctr) {
  return ((double)(orc_rng_u64(seed, ctr) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
double orc_rng_gauss(uint64_t seed, uint64_t ctr) {
  double u1 = rng_uniform(seed, 2 * ctr);
  double u2 = rng_uniform(seed, 2 * ctr + 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static void m3_mul(cpx* r, const cpx* a, const cpx* b) { su3_mul(r, a, b); }

/* exp(X) for a 3x3 complex matrix by scaling and squaring a Taylor series */
static void m3_exp(cpx* e, const cpx* x) {
  cpx xs[9], term[9], tmp[9];
  const int sq = 6;
  const double sc = 1.0 / 64.0;
  for (int k = 0; k < 9; ++k) { xs[k] = cscale(x[k], sc); e[k] = C(k % 4 == 0 ? 1 : 0, 0); term[k] = e[k]; }
  for (int n = 1; n <= 12; ++n) {
    m3_mul(tmp, term, xs);
    for (int k = 0; k < 9; ++k) { term[k] = cscale(tmp[k], 1.0 / n); e[k] = cadd(e[k], term[k]); }
  }
  for (int s = 0; s < sq; ++s) {
    m3_mul(tmp, e, e);
    for (int k = 0; k < 9; ++k) e[k] = tmp[k];
  }
}

/* Gram-Schmidt of rows 0, 1 then row 2 = conj(row0 x row1) */
static void su3_project(cpx* u) {
  double n0 = 0;
  for (int j = 0; j < 3; ++j) n0 += u[j].re * u[j].re + u[j].im * u[j].im;
  n0 = 1.0 / sqrt(n0);
  for (int j = 0; j < 3; ++j) u[j] = cscale(u[j], n0);
  cpx d = C(0, 0);
  for (int j = 0; j < 3; ++j) d = cadd(d, cmul(cconj(u[j]), u[3 + j]));
  for (int j = 0; j < 3; ++j) u[3 + j] = csub(u[3 + j], cmul(d, u[j]));
  double n1 = 0;
  for (int j = 0; j < 3; ++j) n1 += u[3 + j].re * u[3 + j].re + u[3 + j].im * u[3 + j].im;
  n1 = 1.0 / sqrt(n1);
  for (int j = 0; j < 3; ++j) u[3 + j] = cscale(u[3 + j], n1);
  recon12(u);
}

static void random_link(cpx* u, uint64_t seed, uint64_t ctr0, double eps) {
  double h[8];
  for (int a = 0; a < 8; ++a) h[a] = orc_rng_gauss(seed, ctr0 + a);
  const double r3 = 0.57735026918962576451; /* 1/sqrt(3) */
  cpx H[9];
  H[0] = C(0.5 * (h[2] + h[7] * r3), 0);
  H[4] = C(0.5 * (-h[2] + h[7] * r3), 0);
  H[8] = C(-h[7] * r3, 0);
  H[1] = C(0.5 * h[0], -0.5 * h[1]);
  H[3] = cconj(H[1]);
  H[2] = C(0.5 * h[3], -0.5 * h[4]);
  H[6] = cconj(H[2]);
  H[5] = C(0.5 * h[5], -0.5 * h[6]);
  H[7] = cconj(H[5]);
  cpx X[9];
  for (int k = 0; k < 9; ++k) X[k] = C(-eps * H[k].im, eps * H[k].re); /* i eps H */
  m3_exp(u, X);
  su3_project(u);
}

void orc_gauge_random(double* gauge, const int L[4], uint64_t seed, double eps) {
  long vh = vh_of(L);
  cpx* g = (cpx*)gauge;
#pragma omp parallel for schedule(static) num_threads(nthreads())
  for (long l = 0; l < 8 * vh; ++l) random_link(g + l * 9, seed, (uint64_t)l * 8, eps);
}

void orc_source_gauss(double* v, long nsites, uint64_t seed) {
  for (long k = 0; k < nsites * 24; ++k) v[k] = orc_rng_gauss(seed, (uint64_t)k);
}

void orc_source_point(double* v, const int L[4], const int c[4], int spin, int colour) {
  long vh = vh_of(L);
  memset(v, 0, sizeof(double) * vh * 2 * 24);
  int p;
  long idx = orc_index(L, c, &p);
  v[((p * vh + idx) * 12 + spin * 3 + colour) * 2] = 1.0;
}
