/*
 * oracle.c -- CPU FP64 restatement of qdk-chemistry's integral-direct J/K
 * path.  TEST INFRASTRUCTURE ONLY: this library is the parity checker for the
 * HIP Fock-build engine.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product never does.
 *
 * What it restates (semantics, not code):
 *   - compute_shellblock_norm      libint2_direct.cpp:43-59
 *   - compute_shellpairs           libint2_direct.cpp:83-151  (same-centre or
 *                                   ||S_ab||_F >= shell_pair_threshold)
 *   - schwarz_integral(use_2norm)  schwarz.cpp:26-78  (Q_ab = sqrt ||(ab|ab)||_F)
 *   - ERI::build_JK                libint2_direct.cpp:256-626  (canonical loop,
 *                                   density-weighted Schwarz test, degeneracy,
 *                                   J/K digestion, thread-local buffers with
 *                                   ordered reduction, symmetrisation)
 *   - ERI::get_gradients           libint2_direct.cpp:646-890  (ordered pairs,
 *                                   dJ = 1/2 sum P12 P34 d(12|34),
 *                                   dK = -1/4 sum P13 P24 d(12|34), xyz-major;
 *                                   spin pairs: P_alpha + P_beta for dJ,
 *                                   -1/2 sum_s Ps13 Ps24 for dK)
 *   - INCORE J/K                   incore_impl.cpp:131-197 (full tensor)
 * The integrals themselves come from Libint2 in the reference (third party,
 * v2.9.0, not vendored).  Here they are Obara-Saika vertical recursion with
 * Head-Gordon-Pople horizontal transfer (Obara & Saika, JCP 84, 3963 (1986);
 * Head-Gordon & Pople, JCP 89, 5777 (1988)), a different published method
 * from the Rys quadrature the GPU engine uses.
 *
 * Parity pins (tests/test_oracle*.py): Boys values vs mpmath fixtures and the
 * erf closed form; (ss|ss) closed form; 8-fold permutation symmetry,
 * translation invariance and the Schwarz inequality on computed quartets;
 * direct J/K == full-tensor J/K; SCF energies vs the reference's fixture
 * energies (water cc-pVDZ, def2-SVP, 6-31G*); gradients vs the reference's
 * h2o.json fixture and central finite differences.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXL 4        /* max shell angular momentum (g) */
#define OR_LT 20         /* max total cartesian L in tables */
#define OR_MAXPRIM 40
#define OR_BOYS_M 32     /* max Boys order held in the Taylor table */
#define OR_BOYS_DT 0.025
#define OR_BOYS_TMAX 45.0
#define OR_BOYS_NG 1802 /* OR_BOYS_TMAX / OR_BOYS_DT + 2 */
#define OR_BOYS_NT 8     /* Taylor terms */

static const double OR_PI = 3.14159265358979323846;

/* ------------------------------------------------------------------------ */
/* cartesian index tables (CCA order: lx descending, then ly descending)    */
/* ------------------------------------------------------------------------ */
#define NCART(l) (((l) + 1) * ((l) + 2) / 2)
#define CART_OFF(l) ((l) * ((l) + 1) * ((l) + 2) / 6)
static inline int ncart(int l) { return NCART(l); }
static inline int cart_off(int l) { return CART_OFF(l); }
/* index of (lx,ly,lz) inside its own l block */
static inline int cart_local(int lx, int ly, int lz) {
  int l = lx + ly + lz;
  (void)ly;
  return ((l - lx) * (l - lx + 1)) / 2 + lz;
}
static inline int cart_global(int lx, int ly, int lz) {
  return cart_off(lx + ly + lz) + cart_local(lx, ly, lz);
}
static int g_cxyz[CART_OFF(OR_LT + 1)][3];

/* ------------------------------------------------------------------------ */
/* real solid harmonics: coefficient rows for l <= OR_MAXL, m = -l..l       */
/* ------------------------------------------------------------------------ */
static double g_sph[OR_MAXL + 1][2 * OR_MAXL + 1][NCART(OR_MAXL)];

static double dfact(int n) { /* n!! with (-1)!! = 1 */
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}
static double fact(int n) {
  double r = 1.0;
  for (int k = 2; k <= n; ++k) r *= k;
  return r;
}
static double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  return fact(n) / (fact(k) * fact(n - k));
}
/* overlap of two cartesian monomials of the same l, each carrying the x^l
 * normalisation (the convention of the PSI4-mode coefficients) */
static double cart_pair_overlap(const int *a, const int *b, int l) {
  int s0 = a[0] + b[0], s1 = a[1] + b[1], s2 = a[2] + b[2];
  if ((s0 & 1) || (s1 & 1) || (s2 & 1)) return 0.0;
  return dfact(s0 - 1) * dfact(s1 - 1) * dfact(s2 - 1) / dfact(2 * l - 1);
}

static void init_spherical(void) {
  memset(g_sph, 0, sizeof(g_sph));
  for (int l = 0; l <= OR_MAXL; ++l) {
    for (int m = -l; m <= l; ++m) {
      int am = abs(m);
      int vvm = (m < 0) ? 1 : 0; /* 2*v_m */
      double *row = g_sph[l][m + l];
      /* Helgaker, Jorgensen & Olsen, eq. 6.4.47-6.4.50 */
      for (int t = 0; t <= (l - am) / 2; ++t)
        for (int u = 0; u <= t; ++u)
          for (int vv = vvm; vv <= am; vv += 2) {
            int sgn_exp = t + (vv - vvm) / 2;
            double c = ((sgn_exp & 1) ? -1.0 : 1.0) * pow(0.25, t) * binom(l, t) *
                       binom(l - t, am + t) * binom(t, u) * binom(am, vv);
            int px = 2 * t + am - 2 * u - vv;
            int py = 2 * u + vv;
            int pz = l - 2 * t - am;
            if (px < 0 || pz < 0) continue;
            row[cart_local(px, py, pz)] += c;
          }
      /* normalise the pure function under x^l-normalised cartesians */
      double nrm = 0.0;
      int n = ncart(l);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          nrm += row[i] * row[j] *
                 cart_pair_overlap(g_cxyz[cart_off(l) + i], g_cxyz[cart_off(l) + j], l);
      double s = 1.0 / sqrt(nrm);
      for (int i = 0; i < n; ++i) row[i] *= s;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Boys function F_m(T): Taylor table + downward recursion, asymptotic tail */
/* ------------------------------------------------------------------------ */
static double g_boys[OR_BOYS_NG][OR_BOYS_M + OR_BOYS_NT + 1];

static long double boys_series(int m, long double T) {
  /* F_m(T) = exp(-T) sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)) */
  long double term = 1.0L / (2 * m + 1), sum = term;
  for (int k = 1; k < 2000; ++k) {
    term *= 2.0L * T / (2 * m + 2 * k + 1);
    sum += term;
    if (term < sum * 1e-21L) break;
  }
  return expl(-T) * sum;
}

static void init_boys(void) {
  for (int g = 0; g < OR_BOYS_NG; ++g) {
    long double T = g * (long double)OR_BOYS_DT;
    int mtop = OR_BOYS_M + OR_BOYS_NT;
    long double F = boys_series(mtop, T);
    g_boys[g][mtop] = (double)F;
    long double eT = expl(-T);
    for (int m = mtop - 1; m >= 0; --m) {
      F = (2.0L * T * F + eT) / (2 * m + 1);
      g_boys[g][m] = (double)F;
    }
  }
}

static void oracle_init(void);

void oracle_boys(int mmax, double T, double *F) {
  oracle_init(); /* the table is built on first use, also without a handle */
  if (T < OR_BOYS_TMAX - 1.0) {
    int g = (int)(T / OR_BOYS_DT + 0.5);
    double d = g * OR_BOYS_DT - T; /* F_m(T) = sum_k F_{m+k}(T0) (T0-T)^k / k! */
    const double *row = g_boys[g];
    double s = 0.0, pk = 1.0;
    for (int k = 0; k < OR_BOYS_NT; ++k) {
      s += row[mmax + k] * pk;
      pk *= d / (k + 1);
    }
    F[mmax] = s;
    double eT = exp(-T);
    for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + eT) / (2 * m + 1);
  } else {
    double eT = exp(-T);
    F[0] = 0.5 * sqrt(OR_PI / T) * erf(sqrt(T));
    for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - eT) / (2.0 * T);
  }
}

static int g_init_done = 0;
static void oracle_init(void) {
  if (g_init_done) return;
#pragma omp critical(oracle_init)
  {
    if (!g_init_done) {
      for (int l = 0; l <= OR_LT; ++l)
        for (int lx = l; lx >= 0; --lx)
          for (int ly = l - lx; ly >= 0; --ly) {
            int lz = l - lx - ly;
            int gi = cart_global(lx, ly, lz);
            g_cxyz[gi][0] = lx;
            g_cxyz[gi][1] = ly;
            g_cxyz[gi][2] = lz;
          }
      init_spherical();
      init_boys();
      g_init_done = 1;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* basis                                                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
  int l, nprim, atom;
  double O[3];
  double a[OR_MAXPRIM], c[OR_MAXPRIM];
} shell_t;

typedef struct {
  int nsh, nbf, natoms, pure;
  shell_t *sh;
  int *bf0;            /* first AO of each shell */
  double *Z;           /* nuclear charges */
  double *R;           /* natoms x 3 */
  double eri_thr, sp_thr;
  double *Q;           /* nsh x nsh Schwarz */
  int *splist_off;     /* nsh+1 offsets into splist */
  int *splist;         /* for each s1, significant s2 <= s1 (ascending) */
  char *sig;           /* nsh x nsh: 1 for a significant pair (both orders) */
  int nthreads;
} oracle_t;

static inline int nfun(const oracle_t *h, int l) { return h->pure ? 2 * l + 1 : ncart(l); }

/* ------------------------------------------------------------------------ */
/* workspace                                                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
  double *V;      /* VRR  [m][e][f]            */
  double *C[4];   /* contracted [e][f] sets     */
  double *H1, *H2, *H3; /* HRR scratch          */
  double *cart;   /* cartesian quartet          */
  double *tmp;
  size_t vsz, csz, hsz, qsz;
} work_t;

static void work_alloc(work_t *w, int lmax_shell, int deriv) {
  int E = 2 * lmax_shell + deriv, F = 2 * lmax_shell + deriv;
  int M = E + F + 1;
  w->vsz = (size_t)M * cart_off(E + 1) * cart_off(F + 1);
  w->csz = (size_t)cart_off(E + 1) * cart_off(F + 1);
  size_t nq1 = ncart(lmax_shell + deriv);
  w->hsz = (size_t)cart_off(E + 1) * (nq1 * nq1) * cart_off(F + 1) + 16;
  w->qsz = nq1 * nq1 * nq1 * nq1;
  w->V = (double *)malloc(sizeof(double) * w->vsz);
  for (int k = 0; k < 4; ++k) w->C[k] = (double *)malloc(sizeof(double) * w->csz);
  w->H1 = (double *)malloc(sizeof(double) * w->hsz);
  w->H2 = (double *)malloc(sizeof(double) * w->hsz);
  w->H3 = (double *)malloc(sizeof(double) * w->hsz);
  w->cart = (double *)malloc(sizeof(double) * w->qsz * 4);
  w->tmp = (double *)malloc(sizeof(double) * w->qsz * 4);
}
static void work_free(work_t *w) {
  free(w->V);
  for (int k = 0; k < 4; ++k) free(w->C[k]);
  free(w->H1);
  free(w->H2);
  free(w->H3);
  free(w->cart);
  free(w->tmp);
}

/* ------------------------------------------------------------------------ */
/* Obara-Saika VRR for one primitive quartet: [e0|f0]^(m)                   */
/*   V[(m*NE + e)*NF + f], e,f global cartesian indices, |e|<=E, |f|<=F     */
/* ------------------------------------------------------------------------ */
static void vrr_prim(double *V, int E, int F, double p, double q, const double *PA,
                     const double *WP, const double *QC, const double *WQ, double pref,
                     double T) {
  const int M = E + F;
  const int NE = cart_off(E + 1), NF = cart_off(F + 1);
  double Fm[OR_LT * 2 + 2];
  oracle_boys(M, T, Fm);
  const double rho = p * q / (p + q);
  const double o2p = 0.5 / p, o2q = 0.5 / q, o2pq = 0.5 / (p + q);
  const double rp = rho / p, rq = rho / q;
#define VV(m, e, f) V[((size_t)(m) * NE + (e)) * NF + (f)]
  for (int m = 0; m <= M; ++m) VV(m, 0, 0) = pref * Fm[m];
  /* build e with f = 0 */
  for (int le = 1; le <= E; ++le)
    for (int ie = cart_off(le); ie < cart_off(le + 1); ++ie) {
      const int *e = g_cxyz[ie];
      int i = e[0] ? 0 : (e[1] ? 1 : 2);
      int em[3] = {e[0], e[1], e[2]};
      em[i]--;
      int e1 = cart_global(em[0], em[1], em[2]);
      int ne = em[i];
      int e2 = -1;
      if (ne > 0) {
        em[i]--;
        e2 = cart_global(em[0], em[1], em[2]);
      }
      for (int m = 0; m <= M - le; ++m) {
        double v = PA[i] * VV(m, e1, 0) + WP[i] * VV(m + 1, e1, 0);
        if (e2 >= 0) v += ne * o2p * (VV(m, e2, 0) - rp * VV(m + 1, e2, 0));
        VV(m, ie, 0) = v;
      }
    }
  /* build f */
  for (int lf = 1; lf <= F; ++lf)
    for (int jf = cart_off(lf); jf < cart_off(lf + 1); ++jf) {
      const int *f = g_cxyz[jf];
      int i = f[0] ? 0 : (f[1] ? 1 : 2);
      int fm[3] = {f[0], f[1], f[2]};
      fm[i]--;
      int f1 = cart_global(fm[0], fm[1], fm[2]);
      int nf = fm[i];
      int f2 = -1;
      if (nf > 0) {
        fm[i]--;
        f2 = cart_global(fm[0], fm[1], fm[2]);
      }
      for (int le = 0; le <= E; ++le)
        for (int ie = cart_off(le); ie < cart_off(le + 1); ++ie) {
          const int *e = g_cxyz[ie];
          int nei = e[i];
          int e1 = -1;
          if (nei > 0) {
            int em[3] = {e[0], e[1], e[2]};
            em[i]--;
            e1 = cart_global(em[0], em[1], em[2]);
          }
          for (int m = 0; m <= M - le - lf; ++m) {
            double v = QC[i] * VV(m, ie, f1) + WQ[i] * VV(m + 1, ie, f1);
            if (f2 >= 0) v += nf * o2q * (VV(m, ie, f2) - rq * VV(m + 1, ie, f2));
            if (e1 >= 0) v += nei * o2pq * VV(m + 1, e1, f1);
            VV(m, ie, jf) = v;
          }
        }
    }
#undef VV
}

/* ------------------------------------------------------------------------ */
/* HRR: contracted [e0|f0] (e in [la,la+lb], f in [lc,lc+ld]) -> cartesian  */
/*   (ab|cd) block out[a][b][c][d] (local cartesian indices)                */
/* ------------------------------------------------------------------------ */
static void hrr_quartet(const double *Cef, int NFtot, int la, int lb, int lc, int ld,
                        const double *AB, const double *CD, double *out, work_t *w) {
  /* bra transfer: X_k[e][b][f]; e global index (|e| in [la, la+lb-k]), b local */
  const int nfr = cart_off(lc + ld + 1); /* f index range [0, nfr) global */
  double *X = w->H1, *Y = w->H2;
  /* level 0 */
  int NEall = cart_off(la + lb + 1);
  for (int e = 0; e < NEall; ++e)
    for (int f = 0; f < nfr; ++f) X[(size_t)e * nfr + f] = (e >= cart_off(la)) ? Cef[(size_t)e * NFtot + f] : 0.0;
  int nbprev = 1; /* number of b components at level k-1 */
  for (int k = 1; k <= lb; ++k) {
    int nb = ncart(k);
    for (int le = la; le <= la + lb - k; ++le)
      for (int e = cart_off(le); e < cart_off(le + 1); ++e)
        for (int ib = 0; ib < nb; ++ib) {
          const int *b = g_cxyz[cart_off(k) + ib];
          int i = b[0] ? 0 : (b[1] ? 1 : 2);
          int bm[3] = {b[0], b[1], b[2]};
          bm[i]--;
          int ibm = cart_local(bm[0], bm[1], bm[2]);
          const int *ev = g_cxyz[e];
          int ep[3] = {ev[0], ev[1], ev[2]};
          ep[i]++;
          int eplus = cart_global(ep[0], ep[1], ep[2]);
          double *dst = Y + ((size_t)e * nb + ib) * nfr;
          const double *s1 = X + ((size_t)eplus * nbprev + ibm) * nfr;
          const double *s2 = X + ((size_t)e * nbprev + ibm) * nfr;
          for (int f = 0; f < nfr; ++f) dst[f] = s1[f] + AB[i] * s2[f];
        }
    double *t = X;
    X = Y;
    Y = t;
    nbprev = nb;
  }
  /* X now holds (a b | f) with a global in l=la block, b local (nb=ncart(lb)) */
  const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
  /* ket transfer for each (a,b) */
  double *K0 = w->tmp; /* [f global][d local] */
  double *K1 = w->H3;
  for (int ia = 0; ia < na; ++ia)
    for (int ib = 0; ib < nb; ++ib) {
      const double *src = X + ((size_t)(cart_off(la) + ia) * nb + ib) * nfr;
      int ndprev = 1;
      for (int f = 0; f < nfr; ++f) K0[f] = src[f];
      double *KX = K0, *KY = K1;
      for (int k = 1; k <= ld; ++k) {
        int ndk = ncart(k);
        for (int lf = lc; lf <= lc + ld - k; ++lf)
          for (int f = cart_off(lf); f < cart_off(lf + 1); ++f)
            for (int id = 0; id < ndk; ++id) {
              const int *d = g_cxyz[cart_off(k) + id];
              int i = d[0] ? 0 : (d[1] ? 1 : 2);
              int dm[3] = {d[0], d[1], d[2]};
              dm[i]--;
              int idm = cart_local(dm[0], dm[1], dm[2]);
              const int *fv = g_cxyz[f];
              int fp[3] = {fv[0], fv[1], fv[2]};
              fp[i]++;
              int fplus = cart_global(fp[0], fp[1], fp[2]);
              KY[(size_t)f * ndk + id] =
                  KX[(size_t)fplus * ndprev + idm] + CD[i] * KX[(size_t)f * ndprev + idm];
            }
        double *t = KX;
        KX = KY;
        KY = t;
        ndprev = ndk;
      }
      for (int ic = 0; ic < nc; ++ic)
        for (int id = 0; id < nd; ++id)
          out[((size_t)(ia * nb + ib) * nc + ic) * nd + id] = KX[(size_t)(cart_off(lc) + ic) * nd + id];
    }
}

/* cartesian (n1c n2c n3c n4c) -> output functions (pure or cart);
 * tmp must hold 2 * n1c*n2c*n3c*n4c doubles */
static void cart_to_out(const oracle_t *h, const double *in, int l1, int l2, int l3, int l4,
                        double *out, double *tmp) {
  int lv[4] = {l1, l2, l3, l4};
  int dims[4], no[4];
  size_t tot = 1;
  for (int k = 0; k < 4; ++k) {
    dims[k] = ncart(lv[k]);
    no[k] = (h->pure && lv[k] >= 2) ? 2 * lv[k] + 1 : dims[k];
    tot *= dims[k];
  }
  double *bufs[2] = {tmp, tmp + tot};
  const double *src = in;
  for (int k = 0; k < 4; ++k) {
    size_t pre = 1, post = 1;
    for (int j = 0; j < k; ++j) pre *= dims[j];
    for (int j = k + 1; j < 4; ++j) post *= dims[j];
    double *dst = bufs[k & 1];
    if (!(h->pure && lv[k] >= 2)) {
      memcpy(dst, src, sizeof(double) * pre * dims[k] * post);
    } else {
      for (size_t a = 0; a < pre; ++a)
        for (int m = 0; m < no[k]; ++m)
          for (size_t b = 0; b < post; ++b) {
            double s = 0.0;
            for (int c = 0; c < dims[k]; ++c) {
              double co = g_sph[lv[k]][m][c];
              if (co != 0.0) s += co * src[(a * dims[k] + c) * post + b];
            }
            dst[(a * no[k] + m) * post + b] = s;
          }
    }
    dims[k] = no[k];
    src = dst;
  }
  memcpy(out, src, sizeof(double) * (size_t)no[0] * no[1] * no[2] * no[3]);
}

/* ------------------------------------------------------------------------ */
/* contracted shell quartet (ab|cd) and its first derivatives               */
/* ------------------------------------------------------------------------ */
static void quartet_cart(const oracle_t *h, int s1, int s2, int s3, int s4, int deriv,
                         work_t *w, double *blocks /* cart blocks */) {
  const shell_t *A = &h->sh[s1], *B = &h->sh[s2], *Cc = &h->sh[s3], *D = &h->sh[s4];
  const int la = A->l, lb = B->l, lc = Cc->l, ld = D->l;
  const int E = la + lb + deriv, F = lc + ld + deriv;
  const int NE = cart_off(E + 1), NF = cart_off(F + 1);
  double AB[3], CD[3];
  for (int i = 0; i < 3; ++i) {
    AB[i] = A->O[i] - B->O[i];
    CD[i] = Cc->O[i] - D->O[i];
  }
  const double AB2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  const double CD2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
  int nsets = deriv ? 4 : 1;
  for (int k = 0; k < nsets; ++k) memset(w->C[k], 0, sizeof(double) * NE * NF);
  const double pi52 = 2.0 * pow(OR_PI, 2.5);
  for (int i1 = 0; i1 < A->nprim; ++i1)
    for (int i2 = 0; i2 < B->nprim; ++i2) {
      double a = A->a[i1], b = B->a[i2], p = a + b;
      double Kab = exp(-a * b / p * AB2) * A->c[i1] * B->c[i2];
      double P[3], PA[3];
      for (int i = 0; i < 3; ++i) {
        P[i] = (a * A->O[i] + b * B->O[i]) / p;
        PA[i] = P[i] - A->O[i];
      }
      for (int i3 = 0; i3 < Cc->nprim; ++i3)
        for (int i4 = 0; i4 < D->nprim; ++i4) {
          double c = Cc->a[i3], d = D->a[i4], q = c + d;
          double Kcd = exp(-c * d / q * CD2) * Cc->c[i3] * D->c[i4];
          double Q[3], QC[3], W[3], WP[3], WQ[3];
          double PQ2 = 0.0;
          for (int i = 0; i < 3; ++i) {
            Q[i] = (c * Cc->O[i] + d * D->O[i]) / q;
            QC[i] = Q[i] - Cc->O[i];
            W[i] = (p * P[i] + q * Q[i]) / (p + q);
            WP[i] = W[i] - P[i];
            WQ[i] = W[i] - Q[i];
            PQ2 += (P[i] - Q[i]) * (P[i] - Q[i]);
          }
          double rho = p * q / (p + q);
          double pref = pi52 / (p * q * sqrt(p + q)) * Kab * Kcd;
          vrr_prim(w->V, E, F, p, q, PA, WP, QC, WQ, pref, rho * PQ2);
          /* m = 0 slice accumulates into contracted sets */
          const double *V0 = w->V;
          double sc[4] = {1.0, 2.0 * a, 2.0 * b, 2.0 * c};
          for (int k = 0; k < nsets; ++k) {
            double *Ck = w->C[k];
            for (size_t x = 0; x < (size_t)NE * NF; ++x) Ck[x] += sc[k] * V0[x];
          }
        }
    }
  const size_t nq = (size_t)ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
  if (!deriv) {
    hrr_quartet(w->C[0], NF, la, lb, lc, ld, AB, CD, blocks, w);
    return;
  }
  /* derivative blocks: d/dA_i, d/dB_i, d/dC_i  -> blocks[(3*center+i)*nq] */
  const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
  size_t nbig = (size_t)ncart(la + 1) * ncart(lb + 1) * ncart(lc + 1) * ncart(ld + 1);
  double *up = (double *)malloc(sizeof(double) * nbig);
  double *dn = (double *)malloc(sizeof(double) * nbig);
  for (int center = 0; center < 3; ++center) {
    int lu[4] = {la, lb, lc, ld}, lm[4] = {la, lb, lc, ld};
    lu[center]++;
    lm[center]--;
    hrr_quartet(w->C[center + 1], NF, lu[0], lu[1], lu[2], lu[3], AB, CD, up, w);
    if (lm[center] >= 0) hrr_quartet(w->C[0], NF, lm[0], lm[1], lm[2], lm[3], AB, CD, dn, w);
    int nu[4] = {ncart(lu[0]), ncart(lu[1]), ncart(lu[2]), ncart(lu[3])};
    int nm[4] = {lm[0] >= 0 ? ncart(lm[0]) : 0, lm[1] >= 0 ? ncart(lm[1]) : 0,
                 lm[2] >= 0 ? ncart(lm[2]) : 0, lm[3] >= 0 ? ncart(lm[3]) : 0};
    int lv[4] = {la, lb, lc, ld};
    for (int i = 0; i < 3; ++i) {
      double *dst = blocks + (size_t)(3 * center + i) * nq;
      for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib)
          for (int ic = 0; ic < nc; ++ic)
            for (int id = 0; id < nd; ++id) {
              int idx[4] = {ia, ib, ic, id};
              int xyz[4][3];
              for (int k = 0; k < 4; ++k) {
                const int *v = g_cxyz[cart_off(lv[k]) + idx[k]];
                xyz[k][0] = v[0];
                xyz[k][1] = v[1];
                xyz[k][2] = v[2];
              }
              int u[4], dmn[4];
              int cnt = xyz[center][i];
              for (int k = 0; k < 4; ++k) {
                u[k] = idx[k];
                dmn[k] = idx[k];
              }
              int t[3] = {xyz[center][0], xyz[center][1], xyz[center][2]};
              t[i]++;
              u[center] = cart_local(t[0], t[1], t[2]);
              double v = up[((size_t)(u[0] * nu[1] + u[1]) * nu[2] + u[2]) * nu[3] + u[3]];
              if (cnt > 0) {
                t[i] -= 2;
                dmn[center] = cart_local(t[0], t[1], t[2]);
                v -= cnt * dn[((size_t)(dmn[0] * nm[1] + dmn[1]) * nm[2] + dmn[2]) * nm[3] + dmn[3]];
              }
              dst[((size_t)(ia * nb + ib) * nc + ic) * nd + id] = v;
            }
    }
  }
  /* translational invariance: d/dD = -(d/dA + d/dB + d/dC) */
  for (int i = 0; i < 3; ++i) {
    double *dst = blocks + (size_t)(9 + i) * nq;
    for (size_t x = 0; x < nq; ++x)
      dst[x] = -(blocks[(size_t)i * nq + x] + blocks[(size_t)(3 + i) * nq + x] +
                 blocks[(size_t)(6 + i) * nq + x]);
  }
  free(up);
  free(dn);
}

/* public: one contracted quartet in output functions (pure/cart); deriv=1
 * gives 12 blocks ordered (A_x,A_y,A_z,B_x,...,D_z) */
static void quartet_out(const oracle_t *h, int s1, int s2, int s3, int s4, int deriv, work_t *w,
                        double *out) {
  const int l1 = h->sh[s1].l, l2 = h->sh[s2].l, l3 = h->sh[s3].l, l4 = h->sh[s4].l;
  size_t nqc = (size_t)ncart(l1) * ncart(l2) * ncart(l3) * ncart(l4);
  size_t nqo = (size_t)nfun(h, l1) * nfun(h, l2) * nfun(h, l3) * nfun(h, l4);
  int nb = deriv ? 12 : 1;
  double *cart = (double *)malloc(sizeof(double) * nqc * nb);
  double *tmp = (double *)malloc(sizeof(double) * nqc * 2);
  quartet_cart(h, s1, s2, s3, s4, deriv, w, cart);
  for (int k = 0; k < nb; ++k) cart_to_out(h, cart + k * nqc, l1, l2, l3, l4, out + k * nqo, tmp);
  free(cart);
  free(tmp);
}

/* ------------------------------------------------------------------------ */
/* one-electron integrals (S, T, V), primitive OS                            */
/* ------------------------------------------------------------------------ */
/* 1D overlap table S[i][j], i<=imax, j<=jmax (without the sqrt(pi/p)exp factor) */
static void os_1d(double PA, double PB, double p, int imax, int jmax, double *S, int ld) {
  double o2p = 0.5 / p;
  S[0] = 1.0;
  for (int i = 1; i <= imax; ++i) S[i * ld] = PA * S[(i - 1) * ld] + (i > 1 ? (i - 1) * o2p * S[(i - 2) * ld] : 0.0);
  for (int j = 1; j <= jmax; ++j)
    for (int i = 0; i <= imax; ++i) {
      double v = PB * S[i * ld + j - 1];
      if (i > 0) v += i * o2p * S[(i - 1) * ld + j - 1];
      if (j > 1) v += (j - 1) * o2p * S[i * ld + j - 2];
      S[i * ld + j] = v;
    }
}

/* cartesian 1e block for pair (s1,s2): op 0 = S, 1 = T, 2 = V */
static void onee_cart(const oracle_t *h, int s1, int s2, int op, double *out) {
  const shell_t *A = &h->sh[s1], *B = &h->sh[s2];
  const int la = A->l, lb = B->l, na = ncart(la), nb = ncart(lb);
  memset(out, 0, sizeof(double) * na * nb);
  double AB[3], AB2 = 0.0;
  for (int i = 0; i < 3; ++i) {
    AB[i] = A->O[i] - B->O[i];
    AB2 += AB[i] * AB[i];
  }
  const int LD = OR_MAXL + 3;
  for (int i1 = 0; i1 < A->nprim; ++i1)
    for (int i2 = 0; i2 < B->nprim; ++i2) {
      double a = A->a[i1], b = B->a[i2], p = a + b;
      double cc = A->c[i1] * B->c[i2];
      double P[3], PA[3], PB[3];
      for (int i = 0; i < 3; ++i) {
        P[i] = (a * A->O[i] + b * B->O[i]) / p;
        PA[i] = P[i] - A->O[i];
        PB[i] = P[i] - B->O[i];
      }
      double Kab = exp(-a * b / p * AB2);
      if (op <= 1) {
        double S[3][(OR_MAXL + 3) * (OR_MAXL + 3)];
        for (int d = 0; d < 3; ++d) os_1d(PA[d], PB[d], p, la + 1, lb + 2, S[d], LD);
        double s00 = pow(OR_PI / p, 1.5) * Kab * cc;
        for (int ia = 0; ia < na; ++ia)
          for (int ib = 0; ib < nb; ++ib) {
            const int *ea = g_cxyz[cart_off(la) + ia], *eb = g_cxyz[cart_off(lb) + ib];
            double v;
            if (op == 0) {
              v = S[0][ea[0] * LD + eb[0]] * S[1][ea[1] * LD + eb[1]] * S[2][ea[2] * LD + eb[2]];
            } else {
              double Tk[3], Sk[3];
              for (int d = 0; d < 3; ++d) {
                int i = ea[d], j = eb[d];
                Sk[d] = S[d][i * LD + j];
                double t = -2.0 * b * (2 * j + 1) * S[d][i * LD + j] + 4.0 * b * b * S[d][i * LD + j + 2];
                if (j >= 2) t += j * (j - 1) * S[d][i * LD + j - 2];
                Tk[d] = -0.5 * t;
              }
              v = Tk[0] * Sk[1] * Sk[2] + Sk[0] * Tk[1] * Sk[2] + Sk[0] * Sk[1] * Tk[2];
            }
            out[ia * nb + ib] += s00 * v;
          }
      } else {
        /* nuclear attraction: VRR on (e|A^(m)|0), e <= la+lb, then HRR */
        const int E = la + lb;
        const int NE = cart_off(E + 1);
        double Vt[(OR_LT + 1) * CART_OFF(2 * OR_MAXL + 1)];
        double acc[CART_OFF(2 * OR_MAXL + 1)];
        memset(acc, 0, sizeof(double) * NE);
        for (int n = 0; n < h->natoms; ++n) {
          const double *Cn = h->R + 3 * n;
          double PC[3], PC2 = 0.0;
          for (int i = 0; i < 3; ++i) {
            PC[i] = P[i] - Cn[i];
            PC2 += PC[i] * PC[i];
          }
          double Fm[OR_LT + 2];
          oracle_boys(E, p * PC2, Fm);
          double pref = -h->Z[n] * 2.0 * OR_PI / p * Kab * cc;
#define VT(m, e) Vt[(m)*NE + (e)]
          for (int m = 0; m <= E; ++m) VT(m, 0) = pref * Fm[m];
          for (int le = 1; le <= E; ++le)
            for (int ie = cart_off(le); ie < cart_off(le + 1); ++ie) {
              const int *e = g_cxyz[ie];
              int i = e[0] ? 0 : (e[1] ? 1 : 2);
              int em[3] = {e[0], e[1], e[2]};
              em[i]--;
              int e1 = cart_global(em[0], em[1], em[2]);
              int ne = em[i];
              int e2 = -1;
              if (ne > 0) {
                em[i]--;
                e2 = cart_global(em[0], em[1], em[2]);
              }
              for (int m = 0; m <= E - le; ++m) {
                double v = PA[i] * VT(m, e1) - PC[i] * VT(m + 1, e1);
                if (e2 >= 0) v += ne * (0.5 / p) * (VT(m, e2) - VT(m + 1, e2));
                VT(m, ie) = v;
              }
            }
#undef VT
          for (int e = 0; e < NE; ++e) acc[e] += Vt[e];
        }
        /* HRR (a, b+1_i) = (a+1_i, b) + AB_i (a, b) */
        double X[CART_OFF(2 * OR_MAXL + 1) * NCART(OR_MAXL)], Y[CART_OFF(2 * OR_MAXL + 1) * NCART(OR_MAXL)];
        for (int e = 0; e < NE; ++e) X[e] = acc[e];
        int nbp = 1;
        for (int k = 1; k <= lb; ++k) {
          int nbk = ncart(k);
          for (int le = la; le <= la + lb - k; ++le)
            for (int e = cart_off(le); e < cart_off(le + 1); ++e)
              for (int ib = 0; ib < nbk; ++ib) {
                const int *bv = g_cxyz[cart_off(k) + ib];
                int i = bv[0] ? 0 : (bv[1] ? 1 : 2);
                int bm[3] = {bv[0], bv[1], bv[2]};
                bm[i]--;
                int ibm = cart_local(bm[0], bm[1], bm[2]);
                const int *ev = g_cxyz[e];
                int ep[3] = {ev[0], ev[1], ev[2]};
                ep[i]++;
                int eplus = cart_global(ep[0], ep[1], ep[2]);
                Y[e * nbk + ib] = X[eplus * nbp + ibm] + AB[i] * X[e * nbp + ibm];
              }
          memcpy(X, Y, sizeof(double) * NE * nbk);
          nbp = nbk;
        }
        for (int ia = 0; ia < na; ++ia)
          for (int ib = 0; ib < nb; ++ib) out[ia * nb + ib] += X[(cart_off(la) + ia) * nb + ib];
      }
    }
}

static void onee_out(const oracle_t *h, int s1, int s2, int op, double *out) {
  const int la = h->sh[s1].l, lb = h->sh[s2].l;
  const int na = ncart(la), nb = ncart(lb);
  double cart[NCART(OR_MAXL) * NCART(OR_MAXL)], t1[NCART(OR_MAXL) * NCART(OR_MAXL)];
  onee_cart(h, s1, s2, op, cart);
  int oa = nfun(h, la), ob = nfun(h, lb);
  int pa = h->pure && la >= 2, pb = h->pure && lb >= 2;
  if (!h->pure || la < 2) oa = na;
  if (!h->pure || lb < 2) ob = nb;
  /* transform index a */
  for (int m = 0; m < oa; ++m)
    for (int ib = 0; ib < nb; ++ib) {
      double s = 0.0;
      if (pa) {
        for (int c = 0; c < na; ++c) s += g_sph[la][m][c] * cart[c * nb + ib];
      } else {
        s = cart[m * nb + ib];
      }
      t1[m * nb + ib] = s;
    }
  for (int m = 0; m < oa; ++m)
    for (int n = 0; n < ob; ++n) {
      double s = 0.0;
      if (pb) {
        for (int c = 0; c < nb; ++c) s += g_sph[lb][n][c] * t1[m * nb + c];
      } else {
        s = t1[m * nb + n];
      }
      out[m * ob + n] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* public API                                                                */
/* ------------------------------------------------------------------------ */
void oracle_set_threads(oracle_t *h, int n) { h->nthreads = n > 0 ? n : 1; }

static int omp_threads_default(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* One-electron matrix (nbf x nbf, row major): op 0 = S, 1 = T, 2 = V */
void oracle_onee(const oracle_t *h, int op, double *M) {
  const int n = h->nbf;
#pragma omp parallel for schedule(dynamic) num_threads(h->nthreads)
  for (int s1 = 0; s1 < h->nsh; ++s1)
    for (int s2 = 0; s2 <= s1; ++s2) {
      double blk[NCART(OR_MAXL) * NCART(OR_MAXL)];
      onee_out(h, s1, s2, op, blk);
      int n1 = nfun(h, h->sh[s1].l), n2 = nfun(h, h->sh[s2].l);
      for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j) {
          M[(size_t)(h->bf0[s1] + i) * n + h->bf0[s2] + j] = blk[i * n2 + j];
          M[(size_t)(h->bf0[s2] + j) * n + h->bf0[s1] + i] = blk[i * n2 + j];
        }
    }
}

static void compute_schwarz_and_pairs(oracle_t *h) {
  const int nsh = h->nsh;
  h->Q = (double *)calloc((size_t)nsh * nsh, sizeof(double));
  char *sig = (char *)calloc((size_t)nsh * nsh, 1);
  int lmax = 0;
  for (int s = 0; s < nsh; ++s)
    if (h->sh[s].l > lmax) lmax = h->sh[s].l;
#pragma omp parallel num_threads(h->nthreads)
  {
    work_t w;
    work_alloc(&w, lmax, 0);
    size_t nmax = (size_t)ncart(lmax) * ncart(lmax);
    double *blk = (double *)malloc(sizeof(double) * nmax * nmax);
#pragma omp for schedule(dynamic)
    for (int s1 = 0; s1 < nsh; ++s1)
      for (int s2 = 0; s2 <= s1; ++s2) {
        /* Schwarz, schwarz.cpp:60-66 : sqrt of the Frobenius norm of (12|12) */
        quartet_out(h, s1, s2, s1, s2, 0, &w, blk);
        int n1 = nfun(h, h->sh[s1].l), n2 = nfun(h, h->sh[s2].l);
        size_t nn = (size_t)n1 * n2 * n1 * n2;
        double fro = 0.0;
        for (size_t x = 0; x < nn; ++x) fro += blk[x] * blk[x];
        h->Q[(size_t)s1 * nsh + s2] = h->Q[(size_t)s2 * nsh + s1] = sqrt(sqrt(fro));
        /* significant pair, libint2_direct.cpp:120-134 */
        const double *O1 = h->sh[s1].O, *O2 = h->sh[s2].O;
        int same = (O1[0] == O2[0] && O1[1] == O2[1] && O1[2] == O2[2]);
        int s = same;
        if (!same) {
          double ov[NCART(OR_MAXL) * NCART(OR_MAXL)];
          onee_out(h, s1, s2, 0, ov);
          double f2 = 0.0;
          for (int x = 0; x < n1 * n2; ++x) f2 += ov[x] * ov[x];
          s = sqrt(f2) >= h->sp_thr;
        }
        sig[(size_t)s1 * nsh + s2] = sig[(size_t)s2 * nsh + s1] = (char)s;
      }
    free(blk);
    work_free(&w);
  }
  h->splist_off = (int *)malloc(sizeof(int) * (nsh + 1));
  int cnt = 0;
  for (int s1 = 0; s1 < nsh; ++s1)
    for (int s2 = 0; s2 <= s1; ++s2) cnt += sig[(size_t)s1 * nsh + s2];
  h->splist = (int *)malloc(sizeof(int) * (cnt > 0 ? cnt : 1));
  cnt = 0;
  for (int s1 = 0; s1 < nsh; ++s1) {
    h->splist_off[s1] = cnt;
    for (int s2 = 0; s2 <= s1; ++s2)
      if (sig[(size_t)s1 * nsh + s2]) h->splist[cnt++] = s2;
  }
  h->splist_off[nsh] = cnt;
  h->sig = sig;
}

/* Create from flat arrays.  prim_off[nsh+1] indexes exps/coefs. */
oracle_t *oracle_create(int nsh, const int *l, const int *atom, const double *centers,
                        const int *prim_off, const double *exps, const double *coefs, int pure,
                        int natoms, const double *Z, const double *R, double eri_thr,
                        double sp_thr) {
  oracle_init();
  oracle_t *h = (oracle_t *)calloc(1, sizeof(oracle_t));
  h->nsh = nsh;
  h->pure = pure;
  h->natoms = natoms;
  h->eri_thr = eri_thr;
  h->sp_thr = sp_thr;
  h->nthreads = omp_threads_default();
  h->sh = (shell_t *)calloc(nsh, sizeof(shell_t));
  h->bf0 = (int *)malloc(sizeof(int) * (nsh + 1));
  int nbf = 0;
  for (int s = 0; s < nsh; ++s) {
    shell_t *sh = &h->sh[s];
    sh->l = l[s];
    sh->atom = atom[s];
    sh->nprim = prim_off[s + 1] - prim_off[s];
    if (sh->l > OR_MAXL || sh->nprim > OR_MAXPRIM) {
      free(h->sh);
      free(h->bf0);
      free(h);
      return NULL;
    }
    for (int i = 0; i < 3; ++i) sh->O[i] = centers[3 * s + i];
    for (int k = 0; k < sh->nprim; ++k) {
      sh->a[k] = exps[prim_off[s] + k];
      sh->c[k] = coefs[prim_off[s] + k];
    }
    h->bf0[s] = nbf;
    nbf += nfun(h, sh->l);
  }
  h->bf0[nsh] = nbf;
  h->nbf = nbf;
  h->Z = (double *)malloc(sizeof(double) * (natoms > 0 ? natoms : 1));
  h->R = (double *)malloc(sizeof(double) * 3 * (natoms > 0 ? natoms : 1));
  for (int a = 0; a < natoms; ++a) {
    h->Z[a] = Z[a];
    for (int i = 0; i < 3; ++i) h->R[3 * a + i] = R[3 * a + i];
  }
  compute_schwarz_and_pairs(h);
  return h;
}

void oracle_destroy(oracle_t *h) {
  if (!h) return;
  free(h->sh);
  free(h->bf0);
  free(h->Z);
  free(h->R);
  free(h->Q);
  free(h->splist_off);
  free(h->splist);
  free(h->sig);
  free(h);
}

int oracle_nbf(const oracle_t *h) { return h->nbf; }
int oracle_num_pairs(const oracle_t *h) { return h->splist_off[h->nsh]; }
void oracle_schwarz(const oracle_t *h, double *Q) {
  memcpy(Q, h->Q, sizeof(double) * h->nsh * h->nsh);
}
/* significant pairs as (s1, s2) with s2 <= s1 in the reference's order */
void oracle_pairs(const oracle_t *h, int *out) {
  int k = 0;
  for (int s1 = 0; s1 < h->nsh; ++s1)
    for (int x = h->splist_off[s1]; x < h->splist_off[s1 + 1]; ++x) {
      out[2 * k] = s1;
      out[2 * k + 1] = h->splist[x];
      ++k;
    }
}

/* single quartet in output functions; deriv: 0 (1 block) or 1 (12 blocks) */
void oracle_quartet(const oracle_t *h, int s1, int s2, int s3, int s4, int deriv, double *out) {
  int lmax = 0;
  for (int s = 0; s < h->nsh; ++s)
    if (h->sh[s].l > lmax) lmax = h->sh[s].l;
  work_t w;
  work_alloc(&w, lmax, deriv);
  quartet_out(h, s1, s2, s3, s4, deriv, &w, out);
  work_free(&w);
}

/* shell-block infinity norms, libint2_direct.cpp:43-59 */
static void shellblock_norm(const oracle_t *h, const double *P, double *nrm) {
  const int n = h->nbf;
  for (int s1 = 0; s1 < h->nsh; ++s1)
    for (int s2 = 0; s2 < h->nsh; ++s2) {
      double mx = 0.0;
      for (int i = h->bf0[s1]; i < h->bf0[s1 + 1]; ++i)
        for (int j = h->bf0[s2]; j < h->bf0[s2 + 1]; ++j) {
          double v = fabs(P[(size_t)i * n + j]);
          if (v > mx || v != v) mx = (v != v) ? v : mx > v ? mx : v;
        }
      nrm[(size_t)s1 * h->nsh + s2] = mx;
    }
}

static inline double max6(double a, double b, double c, double d, double e, double f) {
  double m = a;
  if (b > m) m = b;
  if (c > m) m = c;
  if (d > m) m = d;
  if (e > m) m = e;
  if (f > m) m = f;
  return m;
}

/*
 * build_JK restatement (libint2_direct.cpp:256-626).  P is ndm x nbf x nbf.
 * Returns 0, or -1 for a non-finite density (libint2_direct.cpp:274-277).
 * stats (optional, length >= 4): [0] surviving shell quartets,
 * [1] primitive quartets of those, [2] function quartets, [3] loop quartets.
 */
static int g_row_stride = 1, g_row_offset = 0;
/* Restrict build_JK to bra shells s1 with s1 % stride == offset (a bounded
 * sample of the same loop, for CPU timing); stride 1 = the full build. */
void oracle_set_row_sample(int stride, int offset) {
  g_row_stride = stride > 0 ? stride : 1;
  g_row_offset = offset;
}

int oracle_build_jk(const oracle_t *h, const double *P, double *J, double *K, double alpha,
                    double beta, int ndm, double *stats) {
  const int nsh = h->nsh, n = h->nbf;
  const size_t nn = (size_t)n * n, msz = (size_t)ndm * nn;
  double *Pn = (double *)malloc(sizeof(double) * nsh * nsh);
  shellblock_norm(h, P, Pn);
  double pmax = 0.0;
  for (size_t x = 0; x < (size_t)nsh * nsh; ++x) {
    if (!isfinite(Pn[x])) {
      free(Pn);
      return -1;
    }
    if (Pn[x] > pmax) pmax = Pn[x];
  }
  const double kscale = alpha + beta;
  const int nthreads = h->nthreads;
  double *Jl = J ? (double *)calloc((size_t)nthreads * msz, sizeof(double)) : NULL;
  double *Kl = K ? (double *)calloc((size_t)nthreads * msz, sizeof(double)) : NULL;
  int lmax = 0;
  for (int s = 0; s < nsh; ++s)
    if (h->sh[s].l > lmax) lmax = h->sh[s].l;
  double st[4] = {0, 0, 0, 0};
#pragma omp parallel num_threads(nthreads) reduction(+ : st[:4])
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    work_t w;
    work_alloc(&w, lmax, 0);
    size_t nmax = (size_t)nfun(h, lmax);
    double *buf = (double *)malloc(sizeof(double) * nmax * nmax * nmax * nmax * 2 + 64);
    double *Jt = J ? Jl + (size_t)tid * msz : NULL;
    double *Kt = K ? Kl + (size_t)tid * msz : NULL;
    size_t s1234 = 0;
    for (int s1 = 0; s1 < nsh; ++s1) {
      if (s1 % g_row_stride != g_row_offset) continue;
      const int bf1 = h->bf0[s1], n1 = h->bf0[s1 + 1] - bf1;
      for (int x12 = h->splist_off[s1]; x12 < h->splist_off[s1 + 1]; ++x12) {
        const int s2 = h->splist[x12];
        const int bf2 = h->bf0[s2], n2 = h->bf0[s2 + 1] - bf2;
        const double P12 = Pn[(size_t)s1 * nsh + s2];
        for (int s3 = 0; s3 <= s1; ++s3) {
          const int bf3 = h->bf0[s3], n3 = h->bf0[s3 + 1] - bf3;
          const double P13 = Pn[(size_t)s1 * nsh + s3], P23 = Pn[(size_t)s2 * nsh + s3];
          const int s4max = (s1 == s3) ? s2 : s3;
          for (int x34 = h->splist_off[s3]; x34 < h->splist_off[s3 + 1]; ++x34) {
            const int s4 = h->splist[x34];
            if (s4 > s4max) break;
            if ((int)((s1234++) % (size_t)nthreads) != tid) continue;
            st[3] += 1.0;
            const double P14 = Pn[(size_t)s1 * nsh + s4], P24 = Pn[(size_t)s2 * nsh + s4],
                         P34 = Pn[(size_t)s3 * nsh + s4];
            const double pn = max6(P12, P13, P23, P14, P24, P34);
            if (pn * h->Q[(size_t)s1 * nsh + s2] * h->Q[(size_t)s3 * nsh + s4] < h->eri_thr) continue;
            const int bf4 = h->bf0[s4], n4 = h->bf0[s4 + 1] - bf4;
            const double deg = (s1 == s2 ? 1.0 : 2.0) * (s3 == s4 ? 1.0 : 2.0) *
                               ((s1 == s3) ? (s2 == s4 ? 1.0 : 2.0) : 2.0);
            quartet_out(h, s1, s2, s3, s4, 0, &w, buf);
            st[0] += 1.0;
            st[1] += (double)h->sh[s1].nprim * h->sh[s2].nprim * h->sh[s3].nprim * h->sh[s4].nprim;
            st[2] += (double)n1 * n2 * n3 * n4;
            for (int idm = 0; idm < ndm; ++idm) {
              const double *Pc = P + idm * nn;
              if (J) {
                double *Jc = Jt + idm * nn;
                for (int i = 0, ijkl = 0; i < n1; ++i)
                  for (int j = 0; j < n2; ++j) {
                    double Jij = 0.0;
                    const double Pij = Pc[(size_t)(bf1 + i) * n + bf2 + j];
                    for (int k = 0; k < n3; ++k)
                      for (int l = 0; l < n4; ++l, ++ijkl) {
                        const double v = buf[ijkl] * deg;
                        Jij += Pc[(size_t)(bf3 + k) * n + bf4 + l] * v;
                        Jc[(size_t)(bf3 + k) * n + bf4 + l] += Pij * v;
                      }
                    Jc[(size_t)(bf1 + i) * n + bf2 + j] += Jij;
                  }
              }
              if (K && kscale != 0.0) {
                double *Kc = Kt + idm * nn;
                for (int i = 0, ijkl = 0; i < n1; ++i)
                  for (int j = 0; j < n2; ++j)
                    for (int k = 0; k < n3; ++k) {
                      double Kik = 0.0, Kjk = 0.0;
                      const double Pik = 0.25 * Pc[(size_t)(bf1 + i) * n + bf3 + k];
                      const double Pjk = 0.25 * Pc[(size_t)(bf2 + j) * n + bf3 + k];
                      for (int l = 0; l < n4; ++l, ++ijkl) {
                        const double v = kscale * buf[ijkl] * deg;
                        Kik += 0.25 * Pc[(size_t)(bf2 + j) * n + bf4 + l] * v;
                        Kjk += 0.25 * Pc[(size_t)(bf1 + i) * n + bf4 + l] * v;
                        Kc[(size_t)(bf1 + i) * n + bf4 + l] += Pjk * v;
                        Kc[(size_t)(bf2 + j) * n + bf4 + l] += Pik * v;
                      }
                      Kc[(size_t)(bf1 + i) * n + bf3 + k] += Kik;
                      Kc[(size_t)(bf2 + j) * n + bf3 + k] += Kjk;
                    }
              }
            }
          }
        }
      }
    }
    free(buf);
    work_free(&w);
  }
  /* ordered reduction + symmetrisation (libint2_direct.cpp:576-625) */
  if (J) {
    memset(J, 0, sizeof(double) * msz);
    for (int t = 0; t < nthreads; ++t)
      for (size_t x = 0; x < msz; ++x) J[x] += Jl[(size_t)t * msz + x];
    for (int idm = 0; idm < ndm; ++idm)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
          double *a = J + idm * nn + (size_t)i * n + j, *b = J + idm * nn + (size_t)j * n + i;
          double v = 0.25 * (*a + *b);
          *a = v;
          *b = v;
        }
    free(Jl);
  }
  if (K) {
    memset(K, 0, sizeof(double) * msz);
    for (int t = 0; t < nthreads; ++t)
      for (size_t x = 0; x < msz; ++x) K[x] += Kl[(size_t)t * msz + x];
    for (int idm = 0; idm < ndm; ++idm)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
          double *a = K + idm * nn + (size_t)i * n + j, *b = K + idm * nn + (size_t)j * n + i;
          double v = 0.5 * (*a + *b);
          *a = v;
          *b = v;
        }
    free(Kl);
  }
  if (stats)
    for (int k = 0; k < 4; ++k) stats[k] = st[k];
  free(Pn);
  return 0;
}

/*
 * Primitive pairs of shells (s1, s2) that the reference's shell-pair data
 * keeps: libint2::ShellPair is built with ln(max_engine_precision),
 * max_engine_precision = DBL_EPSILON / 1e10 (libint2_direct.cpp:29,87,146),
 * and drops a primitive pair whose |c_a c_b| exp(-a b / (a + b) |AB|^2) is
 * below that precision.
 */
static int sig_prim_pairs(const oracle_t *h, int s1, int s2) {
  const shell_t *A = &h->sh[s1], *B = &h->sh[s2];
  double AB2 = 0.0;
  for (int i = 0; i < 3; ++i) AB2 += (A->O[i] - B->O[i]) * (A->O[i] - B->O[i]);
  const double cut = 2.220446049250313e-16 / 1e10;
  int n = 0;
  for (int i = 0; i < A->nprim; ++i)
    for (int j = 0; j < B->nprim; ++j) {
      const double a = A->a[i], b = B->a[j];
      if (!(fabs(A->c[i] * B->c[j] * exp(-a * b / (a + b) * AB2)) < cut)) ++n;
    }
  return n;
}

/*
 * Per-class census of the screened canonical quartet list: for each class
 * (l1,l2,l3,l4) (index ((l1*5+l2)*5+l3)*5+l4), the surviving shell quartets,
 * their primitive quartets (products of shell primitive counts) and, in
 * nprim_sig (optional), their primitive quartets over the significant
 * primitive pairs of sig_prim_pairs.  No integrals are computed.
 */
void oracle_census(const oracle_t *h, const double *P, double *nq, double *nprim, double *nprim_sig) {
  const int nsh = h->nsh;
  double *Pn = (double *)malloc(sizeof(double) * nsh * nsh);
  shellblock_norm(h, P, Pn);
  /* significant primitive pairs per shell pair, only when asked for */
  int *npp = NULL;
  if (nprim_sig) {
    npp = (int *)calloc((size_t)nsh * nsh, sizeof(int));
#pragma omp parallel for schedule(dynamic) num_threads(h->nthreads)
    for (int s1 = 0; s1 < nsh; ++s1)
      for (int x12 = h->splist_off[s1]; x12 < h->splist_off[s1 + 1]; ++x12)
        npp[(size_t)s1 * nsh + h->splist[x12]] = sig_prim_pairs(h, s1, h->splist[x12]);
  }
  const int NC = 625;
  memset(nq, 0, sizeof(double) * NC);
  memset(nprim, 0, sizeof(double) * NC);
  if (nprim_sig) memset(nprim_sig, 0, sizeof(double) * NC);
#pragma omp parallel num_threads(h->nthreads)
  {
    double *lq = (double *)calloc(NC, sizeof(double));
    double *lp = (double *)calloc(NC, sizeof(double));
    double *ls = (double *)calloc(NC, sizeof(double));
#pragma omp for schedule(dynamic)
    for (int s1 = 0; s1 < nsh; ++s1)
      for (int x12 = h->splist_off[s1]; x12 < h->splist_off[s1 + 1]; ++x12) {
        const int s2 = h->splist[x12];
        const double P12 = Pn[(size_t)s1 * nsh + s2];
        for (int s3 = 0; s3 <= s1; ++s3) {
          const double P13 = Pn[(size_t)s1 * nsh + s3], P23 = Pn[(size_t)s2 * nsh + s3];
          const int s4max = (s1 == s3) ? s2 : s3;
          for (int x34 = h->splist_off[s3]; x34 < h->splist_off[s3 + 1]; ++x34) {
            const int s4 = h->splist[x34];
            if (s4 > s4max) break;
            const double pn = max6(P12, P13, P23, Pn[(size_t)s1 * nsh + s4],
                                   Pn[(size_t)s2 * nsh + s4], Pn[(size_t)s3 * nsh + s4]);
            if (pn * h->Q[(size_t)s1 * nsh + s2] * h->Q[(size_t)s3 * nsh + s4] < h->eri_thr) continue;
            int c = ((h->sh[s1].l * 5 + h->sh[s2].l) * 5 + h->sh[s3].l) * 5 + h->sh[s4].l;
            lq[c] += 1.0;
            lp[c] += (double)h->sh[s1].nprim * h->sh[s2].nprim * h->sh[s3].nprim * h->sh[s4].nprim;
            if (npp)
              ls[c] += (double)npp[(size_t)s1 * nsh + s2] * (double)npp[(size_t)s3 * nsh + s4];
          }
        }
      }
#pragma omp critical(oracle_census)
    for (int c = 0; c < NC; ++c) {
      nq[c] += lq[c];
      nprim[c] += lp[c];
      if (nprim_sig) nprim_sig[c] += ls[c];
    }
    free(lq);
    free(lp);
    free(ls);
  }
  free(npp);
  free(Pn);
}

/* INCORE-semantics J/K from the full (unscreened) tensor (incore_impl.cpp:159-197):
 * J_mn = sum_ls P_ls (mn|ls), K_mn = (alpha+beta) sum_ls P_ls (ml|ns).  Small N only. */
/* the full tensor T[i][j][k][l] = (ij|kl) over basis functions, every shell
   quartet computed (no screening) */
static void fill_full(const oracle_t *h, double *T) {
  const int n = h->nbf, nsh = h->nsh;
  int lmax = 0;
  for (int s = 0; s < nsh; ++s)
    if (h->sh[s].l > lmax) lmax = h->sh[s].l;
#pragma omp parallel num_threads(h->nthreads)
  {
    work_t w;
    work_alloc(&w, lmax, 0);
    size_t nmax = (size_t)nfun(h, lmax);
    double *buf = (double *)malloc(sizeof(double) * nmax * nmax * nmax * nmax);
#pragma omp for schedule(dynamic) collapse(2)
    for (int s1 = 0; s1 < nsh; ++s1)
      for (int s2 = 0; s2 < nsh; ++s2)
        for (int s3 = 0; s3 < nsh; ++s3)
          for (int s4 = 0; s4 < nsh; ++s4) {
            quartet_out(h, s1, s2, s3, s4, 0, &w, buf);
            int n1 = h->bf0[s1 + 1] - h->bf0[s1], n2s = h->bf0[s2 + 1] - h->bf0[s2];
            int n3 = h->bf0[s3 + 1] - h->bf0[s3], n4 = h->bf0[s4 + 1] - h->bf0[s4];
            for (int i = 0, x = 0; i < n1; ++i)
              for (int j = 0; j < n2s; ++j)
                for (int k = 0; k < n3; ++k)
                  for (int l = 0; l < n4; ++l, ++x)
                    T[(((size_t)(h->bf0[s1] + i) * n + h->bf0[s2] + j) * n + h->bf0[s3] + k) * n +
                      h->bf0[s4] + l] = buf[x];
          }
    free(buf);
    work_free(&w);
  }
}

/* the dense AO tensor, the input of the quarter transform
   (incore_impl.cpp:242-297 contracts it with C; libint2_direct.cpp:913-1141
   computes the same contraction quartet by quartet) */
int oracle_eri_full(const oracle_t *h, double *T) {
  if (!h || !T) return -1;
  fill_full(h, T);
  return 0;
}

int oracle_jk_incore(const oracle_t *h, const double *P, double *J, double *K, double alpha,
                     double beta) {
  const int n = h->nbf;
  const size_t n2 = (size_t)n * n;
  if ((double)n2 * n2 * 8.0 > 4e9) return -2;
  double *T = (double *)malloc(sizeof(double) * n2 * n2);
  if (!T) return -2;
  fill_full(h, T);
  const double ks = alpha + beta;
#pragma omp parallel for num_threads(h->nthreads)
  for (int m = 0; m < n; ++m)
    for (int nu = 0; nu < n; ++nu) {
      double jv = 0.0, kv = 0.0;
      for (int l = 0; l < n; ++l)
        for (int s = 0; s < n; ++s) {
          jv += P[(size_t)l * n + s] * T[(((size_t)m * n + nu) * n + l) * n + s];
          kv += P[(size_t)l * n + s] * T[(((size_t)m * n + l) * n + nu) * n + s];
        }
      if (J) J[(size_t)m * n + nu] = jv;
      if (K) K[(size_t)m * n + nu] = ks * kv;
    }
  free(T);
  return 0;
}

/*
 * get_gradients restatement (libint2_direct.cpp:646-890) for one density
 * (RHF, ndm = 1).  dJ, dK: 3*natoms, xyz-major (coord = atom + xyz*natoms).
 * mode 0: the reference's ordered-pair loop (num_pairs^2, no pair symmetry);
 * mode 1: canonical quartets with degeneracy weights (same sum, 8x fewer).
 */
/* Restrict oracle_gradients to one ERI class (for locating a faulty class
 * kernel): key = 1000 la + 100 lb + 10 lc + ld of the class with la >= lb,
 * lc >= ld and the (la, lb) pair class not below (lc, ld); -1 = all classes. */
static int g_grad_key = -1;
void oracle_set_grad_class(int key) { g_grad_key = key; }
/* With g_grad_centre = X in 0..2 only the derivative by centre X of each
 * quartet is kept, added to atom(X) and subtracted from atom(D) (the form a
 * kernel that gets D by translation invariance accumulates); -1 = all. */
static int g_grad_centre = -1;
void oracle_set_grad_centre(int x) { g_grad_centre = x; }

static int class_key(int l1, int l2, int l3, int l4) {
  int a = l1 > l2 ? l1 : l2, b = l1 > l2 ? l2 : l1, c = l3 > l4 ? l3 : l4, d = l3 > l4 ? l4 : l3;
  int pab = a * (a + 1) / 2 + b, pcd = c * (c + 1) / 2 + d;
  if (pcd > pab) {
    int t = a; a = c; c = t;
    t = b; b = d; d = t;
  }
  return 1000 * a + 100 * b + 10 * c + d;
}

int oracle_gradients(const oracle_t *h, const double *P, double *dJ, double *dK, double alpha,
                     double beta, int mode, int ndm) {
  const int nsh = h->nsh, n = h->nbf, na = h->natoms;
  const size_t nn = (size_t)n * n;
  double *Pn = (double *)malloc(sizeof(double) * nsh * nsh);
  /* screening with the first density block (libint2_direct.cpp:693) */
  shellblock_norm(h, P, Pn);
  /* Coulomb density: the spin sum of a spin pair (:683-692) */
  double *Ptot = NULL;
  const double *PJ = P;
  if (ndm == 2) {
    Ptot = (double *)malloc(sizeof(double) * nn);
    for (size_t x = 0; x < nn; ++x) Ptot[x] = P[x] + P[nn + x];
    PJ = Ptot;
  }
  /* exchange scale of the ordered sum: -1/4 (one density), -1/2 per spin (:884) */
  const double kfac = ndm == 2 ? -0.5 : -0.25;
  for (size_t x = 0; x < (size_t)nsh * nsh; ++x)
    if (!isfinite(Pn[x])) {
      free(Pn);
      return -1;
    }
  const double ks = alpha + beta;
  int npair = h->splist_off[nsh];
  int *op = (int *)malloc(sizeof(int) * 4 * (npair > 0 ? npair : 1));
  int nop = 0;
  for (int s1 = 0; s1 < nsh; ++s1)
    for (int x = h->splist_off[s1]; x < h->splist_off[s1 + 1]; ++x) {
      int s2 = h->splist[x];
      op[2 * nop] = s1;
      op[2 * nop + 1] = s2;
      ++nop;
      if (mode == 0 && s1 != s2) {
        op[2 * nop] = s2;
        op[2 * nop + 1] = s1;
        ++nop;
      }
    }
  int lmax = 0;
  for (int s = 0; s < nsh; ++s)
    if (h->sh[s].l > lmax) lmax = h->sh[s].l;
  double *gJ = (double *)calloc(3 * na, sizeof(double));
  double *gK = (double *)calloc(3 * na, sizeof(double));
  const long long total = (long long)nop * nop;
#pragma omp parallel num_threads(h->nthreads)
  {
    work_t w;
    work_alloc(&w, lmax, 1);
    size_t nmax = (size_t)nfun(h, lmax);
    double *buf = (double *)malloc(sizeof(double) * 12 * nmax * nmax * nmax * nmax);
    double *lJ = (double *)calloc(3 * na, sizeof(double));
    double *lK = (double *)calloc(3 * na, sizeof(double));
#pragma omp for schedule(dynamic, 64)
    for (long long q = 0; q < total; ++q) {
      int p12 = (int)(q / nop), p34 = (int)(q % nop);
      int s1 = op[2 * p12], s2 = op[2 * p12 + 1], s3 = op[2 * p34], s4 = op[2 * p34 + 1];
      double deg = 1.0;
      if (mode == 1) {
        if (p34 > p12) continue;
        deg = (s1 == s2 ? 1.0 : 2.0) * (s3 == s4 ? 1.0 : 2.0) * (p12 == p34 ? 1.0 : 2.0);
      }
      double pn = max6(Pn[(size_t)s1 * nsh + s2], Pn[(size_t)s1 * nsh + s3], Pn[(size_t)s1 * nsh + s4],
                       Pn[(size_t)s2 * nsh + s3], Pn[(size_t)s2 * nsh + s4], Pn[(size_t)s3 * nsh + s4]);
      if (pn * h->Q[(size_t)s1 * nsh + s2] * h->Q[(size_t)s3 * nsh + s4] < h->eri_thr) continue;
      if (g_grad_key >= 0 && class_key(h->sh[s1].l, h->sh[s2].l, h->sh[s3].l, h->sh[s4].l) != g_grad_key) continue;
      quartet_out(h, s1, s2, s3, s4, 1, &w, buf);
      int b1 = h->bf0[s1], b2 = h->bf0[s2], b3 = h->bf0[s3], b4 = h->bf0[s4];
      int n1 = h->bf0[s1 + 1] - b1, n2 = h->bf0[s2 + 1] - b2, n3 = h->bf0[s3 + 1] - b3,
          n4 = h->bf0[s4 + 1] - b4;
      size_t nq = (size_t)n1 * n2 * n3 * n4;
      int atoms[4] = {h->sh[s1].atom, h->sh[s2].atom, h->sh[s3].atom, h->sh[s4].atom};
      for (int d = 0; d < 12; ++d) {
        const double *sh = buf + d * nq;
        double vj = 0.0, vk = 0.0;
        for (int i = 0, x = 0; i < n1; ++i)
          for (int j = 0; j < n2; ++j)
            for (int k = 0; k < n3; ++k)
              for (int l = 0; l < n4; ++l, ++x) {
                vj += PJ[(size_t)(b1 + i) * n + b2 + j] * PJ[(size_t)(b3 + k) * n + b4 + l] * sh[x];
                /* ordered loop: P_ik P_jl; canonical loop: the permutation
                 * average (P_ik P_jl + P_il P_jk) / 2 times the degeneracy;
                 * summed over the spin densities */
                for (int idm = 0; idm < ndm; ++idm) {
                  const double *Pc = P + idm * nn;
                  const double pk =
                      (mode == 0) ? Pc[(size_t)(b1 + i) * n + b3 + k] * Pc[(size_t)(b2 + j) * n + b4 + l]
                                  : 0.5 * (Pc[(size_t)(b1 + i) * n + b3 + k] * Pc[(size_t)(b2 + j) * n + b4 + l] +
                                           Pc[(size_t)(b1 + i) * n + b4 + l] * Pc[(size_t)(b2 + j) * n + b3 + k]);
                  vk += pk * ks * sh[x];
                }
              }
        int coord = atoms[d / 3] + (d % 3) * na;
        if (g_grad_centre >= 0) {
          if (d / 3 != g_grad_centre) continue;
          int cd = atoms[3] + (d % 3) * na;
          lJ[cd] -= 0.5 * deg * vj;
          lK[cd] -= kfac * deg * vk;
        }
        lJ[coord] += 0.5 * deg * vj;
        lK[coord] += kfac * deg * vk;
      }
    }
#pragma omp critical(oracle_grad)
    for (int c = 0; c < 3 * na; ++c) {
      gJ[c] += lJ[c];
      gK[c] += lK[c];
    }
    free(lJ);
    free(lK);
    free(buf);
    work_free(&w);
  }
  if (dJ) memcpy(dJ, gJ, sizeof(double) * 3 * na);
  if (dK) memcpy(dK, gK, sizeof(double) * 3 * na);
  free(gJ);
  free(gK);
  free(op);
  free(Pn);
  free(Ptot);
  return 0;
}

/*
 * Exact rows of the J and K that build_JK returns, for the functions of the
 * shells rows[0..nrows): Jr[f][nbf] and Kr[f][nbf] with f running over those
 * shells' functions in the given order.  Every ordered quartet (ab|cd) with a
 * a row shell, (a, b) and (c, d) significant pairs, is visited once and kept
 * or skipped by the test the reference applies to its canonical
 * representative: max6 * Q(s1, s2) * Q(s3, s4) < threshold with (s1, s2) the
 * pair of larger (first, second) shell index (libint2_direct.cpp:357-402).
 * Then J_ab = sum_cd (ab|cd) P_cd and K_ac = (alpha + beta) sum_bd (ab|cd)
 * P_bd, which is what the canonical loop's degeneracy-weighted, symmetrised
 * accumulation equals (libint2_direct.cpp:434-625).  One density; a cost of
 * about (shells per row) x (significant pairs) quartets per row shell, so the
 * rows of systems far too large for the full oracle stay cheap.
 */
int oracle_jk_rows(const oracle_t *h, const double *P, int nrows, const int *rows, double *Jr, double *Kr,
                   double alpha, double beta) {
  const int nsh = h->nsh, n = h->nbf;
  double *Pn = (double *)malloc(sizeof(double) * nsh * nsh);
  shellblock_norm(h, P, Pn);
  for (size_t x = 0; x < (size_t)nsh * nsh; ++x)
    if (!isfinite(Pn[x])) {
      free(Pn);
      return -1;
    }
  const double ks = alpha + beta;
  int *roff = (int *)malloc(sizeof(int) * (nrows + 1));
  roff[0] = 0;
  for (int r = 0; r < nrows; ++r) roff[r + 1] = roff[r] + (h->bf0[rows[r] + 1] - h->bf0[rows[r]]);
  const size_t nf = (size_t)roff[nrows];
  memset(Jr, 0, sizeof(double) * nf * n);
  memset(Kr, 0, sizeof(double) * nf * n);
  /* ordered significant pairs (c, d) */
  int nop = 0;
  for (int c = 0; c < nsh; ++c)
    for (int d = 0; d < nsh; ++d) nop += h->sig[(size_t)c * nsh + d];
  int *op = (int *)malloc(sizeof(int) * 2 * (nop > 0 ? nop : 1));
  nop = 0;
  for (int c = 0; c < nsh; ++c)
    for (int d = 0; d < nsh; ++d)
      if (h->sig[(size_t)c * nsh + d]) {
        op[2 * nop] = c;
        op[2 * nop + 1] = d;
        ++nop;
      }
  int lmax = 0;
  for (int s = 0; s < nsh; ++s)
    if (h->sh[s].l > lmax) lmax = h->sh[s].l;
  /* work units: (row, b) with (a, b) significant */
  int nunit = 0;
  for (int r = 0; r < nrows; ++r)
    for (int b = 0; b < nsh; ++b) nunit += h->sig[(size_t)rows[r] * nsh + b];
  int *unit = (int *)malloc(sizeof(int) * 2 * (nunit > 0 ? nunit : 1));
  nunit = 0;
  for (int r = 0; r < nrows; ++r)
    for (int b = 0; b < nsh; ++b)
      if (h->sig[(size_t)rows[r] * nsh + b]) {
        unit[2 * nunit] = r;
        unit[2 * nunit + 1] = b;
        ++nunit;
      }
  const int nth = h->nthreads;
  double *Jl = (double *)calloc((size_t)nth * nf * n, sizeof(double));
  double *Kl = (double *)calloc((size_t)nth * nf * n, sizeof(double));
#pragma omp parallel num_threads(nth)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    work_t w;
    work_alloc(&w, lmax, 0);
    size_t nmax = (size_t)nfun(h, lmax);
    double *buf = (double *)malloc(sizeof(double) * nmax * nmax * nmax * nmax * 2 + 64);
    double *Jt = Jl + (size_t)tid * nf * n, *Kt = Kl + (size_t)tid * nf * n;
#pragma omp for schedule(dynamic, 1)
    for (int u = 0; u < nunit; ++u) {
      const int r = unit[2 * u], a = rows[r], b = unit[2 * u + 1];
      const int ba = h->bf0[a], na_ = h->bf0[a + 1] - ba, bb = h->bf0[b], nb_ = h->bf0[b + 1] - bb;
      const int p1 = a > b ? a : b, p2 = a > b ? b : a;
      const double Qab = h->Q[(size_t)a * nsh + b];
      for (int x = 0; x < nop; ++x) {
        const int c = op[2 * x], d = op[2 * x + 1];
        const int q1 = c > d ? c : d, q2 = c > d ? d : c;
        const double Qcd = h->Q[(size_t)c * nsh + d];
        const double pn = max6(Pn[(size_t)a * nsh + b], Pn[(size_t)a * nsh + c], Pn[(size_t)b * nsh + c],
                               Pn[(size_t)a * nsh + d], Pn[(size_t)b * nsh + d], Pn[(size_t)c * nsh + d]);
        const int bra_first = p1 > q1 || (p1 == q1 && p2 >= q2);
        if (pn * (bra_first ? Qab : Qcd) * (bra_first ? Qcd : Qab) < h->eri_thr) continue;
        const int bc = h->bf0[c], nc_ = h->bf0[c + 1] - bc, bd = h->bf0[d], nd_ = h->bf0[d + 1] - bd;
        quartet_out(h, a, b, c, d, 0, &w, buf);
        for (int i = 0, ijkl = 0; i < na_; ++i) {
          double *Jrow = Jt + (size_t)(roff[r] + i) * n;
          double *Krow = Kt + (size_t)(roff[r] + i) * n;
          for (int j = 0; j < nb_; ++j) {
            double jab = 0.0;
            for (int k = 0; k < nc_; ++k) {
              double kac = 0.0;
              for (int l = 0; l < nd_; ++l, ++ijkl) {
                const double v = buf[ijkl];
                jab += v * P[(size_t)(bc + k) * n + bd + l];
                kac += v * P[(size_t)(bb + j) * n + bd + l];
              }
              Krow[bc + k] += ks * kac;
            }
            Jrow[bb + j] += jab;
          }
        }
      }
    }
    free(buf);
    work_free(&w);
  }
  for (int t = 0; t < nth; ++t)
    for (size_t x = 0; x < nf * n; ++x) {
      Jr[x] += Jl[(size_t)t * nf * n + x];
      Kr[x] += Kl[(size_t)t * nf * n + x];
    }
  free(Jl);
  free(Kl);
  free(unit);
  free(op);
  free(roff);
  free(Pn);
  return 0;
}

/* expose the spherical coefficient table (for tests): row m of shell l */
void oracle_sph_row(int l, int m, double *row) {
  oracle_init();
  memcpy(row, g_sph[l][m + l], sizeof(double) * ncart(l));
}
