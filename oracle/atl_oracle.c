/*
 * atl_oracle.c -- CPU oracle for the FP64 level-3 path (TEST INFRASTRUCTURE).
 *
 * A plain-C restatement of what ATLAS 3.11.42's own testers do: the LCG
 * matrix generator, the norms, the residual formulas of l3blastst, gemmtst,
 * lutst, llttst and slvtst, and textbook reference routines (refgemm loop
 * order, unblocked right-looking LU with first-index argmax, unblocked
 * Cholesky, substitution TRSM).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, as the checker.  The GPU library never
 * links it.  Citations are basefile:line in the ATLAS source tree.
 *
 * Pinning (what independently checks each part) is listed in the test files
 * tests/test_oracle_*.py; every function below is covered there.
 */
#include "atl_oracle.h"
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <pthread.h>

enum { NoTrans = 111, Trans = 112, ConjTrans = 113, Upper = 121, Lower = 122,
       NonUnit = 131, Unit = 132, Left = 141, Right = 142 };

/* ------------------------------------------------------------------------
 * LCG.  AtlasBase/Clint/atlas-tst.base:5353-5620 (gen src/testing/ATL_rand.c)
 * 31-bit numbers are carried as a 16-bit low half and a 15-bit high half;
 * products wrap in 32-bit signed arithmetic and a negative intermediate is
 * folded back by adding 2^31 (ATL_lmul).  The wrap is done on unsigned
 * values here so that the C code is well defined.
 * ------------------------------------------------------------------------ */
static int wmul(int a, int b) { return (int)((unsigned)a * (unsigned)b); }
static int wadd(int a, int b) { return (int)((unsigned)a + (unsigned)b); }
static int fold31(int k) { return k < 0 ? (k + (1 << 30)) + (1 << 30) : k; }

/* ATL_lmul: I = K*J mod 2^31 on (16,15)-bit halves (atlas-tst.base:5430+) */
static void lmul(const int *K, const int *J, int *I)
{
   int kt = fold31(wmul(K[0], J[0]));
   int lt;
   I[0] = kt - ((kt >> 16) << 16);
   lt = fold31(wadd(wmul(K[0], J[1]), wmul(K[1], J[0])));
   kt = fold31(wadd(kt >> 16, lt));
   I[1] = kt - ((kt >> 15) << 15);
}

/* ATL_ladd: I = J + K without carry out of bit 31 (atlas-tst.base:5390+) */
static void ladd(const int *J, const int *K, int *I)
{
   int t0 = K[0] + J[0], t1;
   t1 = t0 >> 16;
   I[0] = t0 - (t1 << 16);
   t0 = t1 + K[1] + J[1];
   I[1] = t0 - ((t0 >> 15) << 15);
}

static int g_ias[2], g_ics[2], g_irand[2];

/* ATL_srand: mult=(20077,16838) iadd=(12345,0); jseed=(iseed, iseed>>16)
 * with the low half NOT masked; state = one jump from jseed
 * (atlas-tst.base:5598-5615) */
void orc_srand(int iseed)
{
   const int mult[2] = {20077, 16838}, iadd[2] = {12345, 0};
   int jseed[2], j[2];
   jseed[0] = iseed;
   jseed[1] = iseed >> 16;
   lmul(jseed, mult, j);          /* ATL_xjumpm with JUMPM=1 */
   ladd(j, iadd, g_irand);
   g_ias[0] = mult[0]; g_ias[1] = mult[1];
   g_ics[0] = iadd[0]; g_ics[1] = iadd[1];
}

/* ATL_rand: return the current state, then advance it (ATL_setran opt 3) */
int orc_rand(void)
{
   int cur[2], j[2];
   cur[0] = g_irand[0]; cur[1] = g_irand[1];
   lmul(g_irand, g_ias, j);
   ladd(j, g_ics, g_irand);
   return cur[0] + (cur[1] << 16);
}

/* dumb_rand() = 0.5 - ATL_rand()/2^31 (gen include/atlas_tst.h:17-18) */
double orc_dumb_rand(void)
{
   return 0.5 - ((double)orc_rand()) / 2147483648.0;
}

void orc_rand_words(int iseed, int n, int *out)
{
   int i;
   orc_srand(iseed);
   for (i = 0; i < n; i++) out[i] = orc_rand();
}

#define FILLCONST (-2560000000.0)

/* ATL_gegen: seed, fill lda gap with FILLCONST, then column-by-column
 * (atlas-tst.base:2949-3005, gen src/testing/ATL_gegen.c) */
void orc_gegen(int M, int N, double *A, int lda, int seed)
{
   int i, j;
   orc_srand(seed);
   if (lda > M)
      for (j = 0; j < N; j++)
         for (i = M; i < lda; i++) A[(size_t)j * lda + i] = FILLCONST;
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++) A[(size_t)j * lda + i] = orc_dumb_rand();
}

/* ATL_gechkgap: count lda-gap entries that changed (atlas-tst.base:2953-2990) */
int orc_gechkgap(int M, int N, const double *A, int lda)
{
   int i, j, n = 0;
   for (j = 0; j < N; j++)
      for (i = M; i < lda; i++)
         if (A[(size_t)j * lda + i] != FILLCONST) n++;
   return n;
}

/* trddom: scale the strict triangle by 1/N and push the diagonal away from
 * zero by one (AtlasBase/Antoine/atlbinblas.base:11727-11790,
 * gen bin/l3blastst.c:539-595) */
void orc_trddom(int uplo, int N, double *A, int lda)
{
   int i, j;
   double alpha;
   if (N <= 0) return;
   alpha = 1.0 / (double)N;
   if (uplo == Upper)
   {
      for (j = 0; j < N; j++)
      {
         double *a = A + (size_t)j * lda;
         for (i = 0; i < j; i++) a[i] *= alpha;
         if (a[j] >= 0.0) a[j] += 1.0; else a[j] -= 1.0;
      }
   }
   else
   {
      for (j = N - 1; j >= 0; j--)
      {
         double *a = A + (size_t)j * lda;
         if (a[j] >= 0.0) a[j] += 1.0; else a[j] -= 1.0;
         for (i = j + 1; i < N; i++) a[i] *= alpha;
      }
   }
}

/* lltgen: gegen, zero the unused triangle (PADVAL = 0), then diagonal
 * a_jj <- |5.3*a_jj| + lltnrm1(A) (AtlasBase/Clint/atlas-tlp.base:1320-1345,
 * gen bin/llttst.c:225-259) */
void orc_lltgen(int uplo, int N, double *A, int lda, int seed)
{
   int i, j;
   double t0, t1;
   orc_gegen(N, N, A, lda, seed);
   for (j = 0; j < N; j++)
   {
      if (uplo == Upper)
         for (i = j + 1; i < N; i++) A[(size_t)j * lda + i] = 0.0;
      else
         for (i = 0; i < j; i++) A[(size_t)j * lda + i] = 0.0;
   }
   t1 = orc_lltnrm1(uplo, N, A, lda);
   for (j = 0; j < N; j++)
   {
      t0 = A[(size_t)j * lda + j] * 5.3;
      if (t0 < 0.0) t0 = -t0;
      A[(size_t)j * lda + j] = t0 + t1;
   }
}

/* ------------------------------------------------------------------------
 * scalars and norms
 * ------------------------------------------------------------------------ */

/* ATL_epsilon: halve until 1+f == 1 (atlas-tst.base:3052-3075) */
double orc_epsilon(void)
{
   double eps;
   volatile double maxval, f1 = 0.5;
   do
   {
      eps = f1;
      f1 *= 0.5;
      maxval = 1.0 + f1;
   } while (maxval != 1.0);
   return eps;
}

/* ATL_genrm1: max column abs-sum, NaN propagates (atlas-tst.base:2877+) */
double orc_genrm1(int M, int N, const double *A, int lda)
{
   int i, j;
   double max = 0.0, t0;
   for (j = 0; j < N; j++)
   {
      t0 = 0.0;
      for (i = 0; i < M; i++) t0 += fabs(A[(size_t)j * lda + i]);
      if (t0 != t0) return t0;
      if (t0 > max) max = t0;
   }
   return max;
}

/* ATL_gediffnrm1: 1-norm of A-B (atlas-tst.base:2333+) */
double orc_gediffnrm1(int M, int N, const double *A, int lda,
                      const double *B, int ldb)
{
   int i, j;
   double max = 0.0, t0;
   for (j = 0; j < N; j++)
   {
      t0 = 0.0;
      for (i = 0; i < M; i++)
         t0 += fabs(A[(size_t)j * lda + i] - B[(size_t)j * ldb + i]);
      if (t0 != t0) return t0;
      if (t0 > max) max = t0;
   }
   return max;
}

/* ATL_trnrm1: triangular 1-norm, unit diagonal counted as 1
 * (gen src/testing/ATL_trnrm1.c, from atlas-tst.base) */
double orc_trnrm1(int uplo, int diag, int N, const double *A, int lda)
{
   int i, j;
   double max = 0.0, t0;
   for (j = 0; j < N; j++)
   {
      const double *a = A + (size_t)j * lda;
      t0 = 0.0;
      if (uplo == Upper)
         for (i = 0; i < j; i++) t0 += fabs(a[i]);
      else
         for (i = j + 1; i < N; i++) t0 += fabs(a[i]);
      if (diag == NonUnit) t0 += fabs(a[j]);
      if (t0 != t0) return t0;
      if (diag == Unit) t0 += 1.0;
      if (t0 > max) max = t0;
   }
   return max;
}

/* lltnrm1: 1-norm of a symmetric matrix from one stored triangle
 * (atlas-tlp.base:1450+, gen bin/llttst.c:166-191) */
double orc_lltnrm1(int uplo, int N, const double *A, int lda)
{
   int i, k;
   double tmp, nrm = 0.0;
   for (i = 0; i < N; i++)
   {
      tmp = 0.0;
      if (uplo == Upper)
      {
         for (k = 0; k < i; k++) tmp += fabs(A[(size_t)i * lda + k]);
         for (k = i; k < N; k++) tmp += fabs(A[(size_t)k * lda + i]);
      }
      else
      {
         for (k = 0; k < i; k++) tmp += fabs(A[(size_t)k * lda + i]);
         for (k = i; k < N; k++) tmp += fabs(A[(size_t)i * lda + k]);
      }
      if (tmp > nrm) nrm = tmp;
   }
   return nrm;
}

double orc_infnrm(int N, const double *x, int incx)
{
   int i;
   double m = 0.0;
   for (i = 0; i < N; i++)
   {
      double t = fabs(x[(size_t)i * incx]);
      if (t > m) m = t;
   }
   return m;
}

/* ------------------------------------------------------------------------
 * textbook routines
 * ------------------------------------------------------------------------ */

static void scal_col(int M, double beta, double *c)
{
   int i;
   if (beta == 0.0) for (i = 0; i < M; i++) c[i] = 0.0;
   else if (beta != 1.0) for (i = 0; i < M; i++) c[i] *= beta;
}

/* ATL_drefgemm + NN/NT/TN/TT loop orders
 * (AtlasBase/Antoine/atlrefblas.base:6801-7100, gen
 * src/blas/reference/level3/ATL_drefgemm{,NN,NT,TN,TT}.c):
 * N-variants accumulate C(:,j) += A(:,l)*(alpha*B(l,j)) after scaling C(:,j)
 * by beta; T-variants form t0 = sum_l A(l,i)*B(l,j) then C = beta*C+alpha*t0.
 * beta==0 overwrites C without reading it. */
void orc_refgemm(int ta, int tb, int M, int N, int K, double alpha,
                 const double *A, int lda, const double *B, int ldb,
                 double beta, double *C, int ldc)
{
   int i, j, l;
   if (M == 0 || N == 0 || ((alpha == 0.0 || K == 0) && beta == 1.0)) return;
   if (alpha == 0.0)
   {
      for (j = 0; j < N; j++) scal_col(M, beta, C + (size_t)j * ldc);
      return;
   }
   for (j = 0; j < N; j++)
   {
      double *c = C + (size_t)j * ldc;
      if (ta == NoTrans)
      {
         scal_col(M, beta, c);
         for (l = 0; l < K; l++)
         {
            const double *a = A + (size_t)l * lda;
            double t0 = alpha * (tb == NoTrans ? B[(size_t)j * ldb + l]
                                               : B[(size_t)l * ldb + j]);
            for (i = 0; i < M; i++) c[i] += a[i] * t0;
         }
      }
      else
      {
         for (i = 0; i < M; i++)
         {
            const double *a = A + (size_t)i * lda;
            double t0 = 0.0;
            if (tb == NoTrans)
               for (l = 0; l < K; l++) t0 += a[l] * B[(size_t)j * ldb + l];
            else
               for (l = 0; l < K; l++) t0 += a[l] * B[(size_t)l * ldb + j];
            if (beta == 0.0) c[i] = 0.0;
            else if (beta != 1.0) c[i] *= beta;
            c[i] += alpha * t0;
         }
      }
   }
}

/* SYRK, one triangle: C <- alpha*op(A)*op(A)^T + beta*C.  Textbook form of
 * ATL_drefsyrk (atlrefblas.base, refsyrk) */
void orc_refsyrk(int uplo, int trans, int N, int K, double alpha,
                 const double *A, int lda, double beta, double *C, int ldc)
{
   int i, j, l;
   for (j = 0; j < N; j++)
   {
      int i0 = (uplo == Upper) ? 0 : j, i1 = (uplo == Upper) ? j + 1 : N;
      for (i = i0; i < i1; i++)
      {
         double t0 = 0.0, *c = C + (size_t)j * ldc + i;
         for (l = 0; l < K; l++)
         {
            double ail, ajl;
            if (trans == NoTrans)
            {
               ail = A[(size_t)l * lda + i];
               ajl = A[(size_t)l * lda + j];
            }
            else
            {
               ail = A[(size_t)i * lda + l];
               ajl = A[(size_t)j * lda + l];
            }
            t0 += ail * ajl;
         }
         if (beta == 0.0) *c = alpha * t0;
         else *c = beta * (*c) + alpha * t0;
      }
   }
}

/* element (i,j) of op(A) */
static double opA(int ta, const double *A, int lda, int i, int j)
{
   return (ta == NoTrans) ? A[(size_t)j * lda + i] : A[(size_t)i * lda + j];
}

/* TRSM by substitution: solves op(A)X = alpha*B (Left) or X op(A) = alpha*B
 * (Right); the Unit diagonal is never read.  Textbook form of the trusted
 * ATL_dtrsm_APR result (AtlasBase/Antoine/atllv3blas.base:399+). */
void orc_reftrsm(int side, int uplo, int ta, int diag, int M, int N,
                 double alpha, const double *A, int lda, double *B, int ldb)
{
   int i, j, k;
   /* effective lower after applying op() */
   const int lower = ((uplo == Lower) == (ta == NoTrans));
   if (M == 0 || N == 0) return;
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++) B[(size_t)j * ldb + i] *= alpha;
   if (side == Left)
   {
      for (j = 0; j < N; j++)
      {
         double *b = B + (size_t)j * ldb;
         if (lower)
            for (i = 0; i < M; i++)
            {
               double s = b[i];
               for (k = 0; k < i; k++) s -= opA(ta, A, lda, i, k) * b[k];
               b[i] = (diag == Unit) ? s : s / opA(ta, A, lda, i, i);
            }
         else
            for (i = M - 1; i >= 0; i--)
            {
               double s = b[i];
               for (k = i + 1; k < M; k++) s -= opA(ta, A, lda, i, k) * b[k];
               b[i] = (diag == Unit) ? s : s / opA(ta, A, lda, i, i);
            }
      }
   }
   else
   {  /* X op(A) = B: column j of X depends on columns k of X with
         op(A)(k,j) != 0, k != j */
      if (lower)
         for (j = N - 1; j >= 0; j--)
         {
            for (k = j + 1; k < N; k++)
            {
               double akj = opA(ta, A, lda, k, j);
               for (i = 0; i < M; i++)
                  B[(size_t)j * ldb + i] -= B[(size_t)k * ldb + i] * akj;
            }
            if (diag == NonUnit)
            {
               double d = opA(ta, A, lda, j, j);
               for (i = 0; i < M; i++) B[(size_t)j * ldb + i] /= d;
            }
         }
      else
         for (j = 0; j < N; j++)
         {
            for (k = 0; k < j; k++)
            {
               double akj = opA(ta, A, lda, k, j);
               for (i = 0; i < M; i++)
                  B[(size_t)j * ldb + i] -= B[(size_t)k * ldb + i] * akj;
            }
            if (diag == NonUnit)
            {
               double d = opA(ta, A, lda, j, j);
               for (i = 0; i < M; i++) B[(size_t)j * ldb + i] /= d;
            }
         }
   }
}

/* cblas_idamax semantics: first index of max |x|, strict '>' against a
 * running max that starts at zero, so NaN is never selected and an all-zero
 * vector gives 0 (AtlasBase/Antoine/atlrefblas.base:753+, gen
 * src/blas/reference/level1/ATL_idrefamax.c) */
int orc_idamax(int N, const double *x, int incx)
{
   int i, imax = 0;
   double m = 0.0;
   if (N <= 0) return 0;
   for (i = 0; i < N; i++)
   {
      double t = fabs(x[(size_t)i * incx]);
      if (t > m) { m = t; imax = i; }
   }
   return imax;
}

#define ORC_SAFMIN DBL_MIN   /* dlamch('S') for IEEE double */

/* Unblocked right-looking LU with partial pivoting, ATLAS getf2 semantics
 * (AtlasBase/Clint/atlas-lp.base:5204-5390, gen src/lapack/ATL_getf2.c):
 * ipiv[j] = absolute 0-based pivot row; the whole panel row is swapped;
 * column scaled by the reciprocal pivot when |p| >= SAFMIN, else divided
 * element by element; a zero pivot sets info (first, 1-based) and the
 * factorization continues.  ATLAS's own tiny-pivot branch divides only a
 * single element (gen ATL_getf2.c:53-55); this oracle divides the column. */
int orc_getf2(int M, int N, double *A, int lda, int *ipiv)
{
   const int MN = M < N ? M : N;
   int i, j, k, info = 0;
   for (j = 0; j < MN; j++)
   {
      double *ac = A + (size_t)j * lda;
      int ip = j + orc_idamax(M - j, ac + j, 1);
      double p = ac[ip];
      ipiv[j] = ip;
      if (p != 0.0)
      {
         if (ip != j)
            for (k = 0; k < N; k++)
            {
               double t = A[(size_t)k * lda + j];
               A[(size_t)k * lda + j] = A[(size_t)k * lda + ip];
               A[(size_t)k * lda + ip] = t;
            }
         if (fabs(p) >= ORC_SAFMIN)
         {
            double r = 1.0 / p;
            for (i = j + 1; i < M; i++) ac[i] *= r;
         }
         else
            for (i = j + 1; i < M; i++) ac[i] /= p;
      }
      else if (!info)
         info = j + 1;
      /* rank-1 update of the trailing panel */
      for (k = j + 1; k < N; k++)
      {
         double *c = A + (size_t)k * lda;
         double u = c[j];
         for (i = j + 1; i < M; i++) c[i] -= ac[i] * u;
      }
   }
   return info;
}

/* ATL_dlaswp: apply interchanges ipiv[K1..K2) in order (inci=1) or in
 * reverse (inci=-1) to N columns (AtlasBase/Clint/atlas-lp.base:1053-1166) */
void orc_laswp(int N, double *A, int lda, int K1, int K2, const int *ipiv,
               int inci)
{
   int i, j;
   if (K2 <= K1) return;
   for (j = 0; j < N; j++)
   {
      double *a = A + (size_t)j * lda;
      if (inci > 0)
         for (i = K1; i < K2; i++)
         {
            int p = ipiv[i];
            double t = a[i]; a[i] = a[p]; a[p] = t;
         }
      else
         for (i = K2 - 1; i >= K1; i--)
         {
            int p = ipiv[i];
            double t = a[i]; a[i] = a[p]; a[p] = t;
         }
   }
}

/* Unblocked Cholesky, ATL_potrfL/U semantics (atlas-lp.base:675-968):
 * info = j+1 for the first non-positive (or NaN) pivot, and stop there.
 * The diagonal is sqrt(d) and the column is scaled by its reciprocal. */
int orc_potf2(int uplo, int N, double *A, int lda)
{
   int i, j, k;
   for (j = 0; j < N; j++)
   {
      double d = A[(size_t)j * lda + j];
      if (uplo == Lower)
         for (k = 0; k < j; k++)
            d -= A[(size_t)k * lda + j] * A[(size_t)k * lda + j];
      else
         for (k = 0; k < j; k++)
            d -= A[(size_t)j * lda + k] * A[(size_t)j * lda + k];
      if (!(d > 0.0)) return j + 1;
      d = sqrt(d);
      A[(size_t)j * lda + j] = d;
      {
         double r = 1.0 / d;
         for (i = j + 1; i < N; i++)
         {
            double s;
            if (uplo == Lower)
            {
               s = A[(size_t)j * lda + i];
               for (k = 0; k < j; k++)
                  s -= A[(size_t)k * lda + i] * A[(size_t)k * lda + j];
               A[(size_t)j * lda + i] = s * r;
            }
            else
            {
               s = A[(size_t)i * lda + j];
               for (k = 0; k < j; k++)
                  s -= A[(size_t)i * lda + k] * A[(size_t)j * lda + k];
               A[(size_t)i * lda + j] = s * r;
            }
         }
      }
   }
   return 0;
}

/* ATL_dgetrs column-major (atlas-lp.base:1631-1720): NoTrans: laswp(+1),
 * L unit solve, U solve.  Trans: U^T solve, L^T unit solve, laswp(-1). */
void orc_getrs(int trans, int N, int NRHS, const double *A, int lda,
               const int *ipiv, double *B, int ldb)
{
   if (!N || !NRHS) return;
   if (trans == NoTrans)
   {
      orc_laswp(NRHS, B, ldb, 0, N, ipiv, 1);
      orc_reftrsm(Left, Lower, NoTrans, Unit, N, NRHS, 1.0, A, lda, B, ldb);
      orc_reftrsm(Left, Upper, NoTrans, NonUnit, N, NRHS, 1.0, A, lda, B, ldb);
   }
   else
   {
      orc_reftrsm(Left, Upper, Trans, NonUnit, N, NRHS, 1.0, A, lda, B, ldb);
      orc_reftrsm(Left, Lower, Trans, Unit, N, NRHS, 1.0, A, lda, B, ldb);
      orc_laswp(NRHS, B, ldb, 0, N, ipiv, -1);
   }
}

/* ATL_dpotrs column-major (atlas-lp.base:969-1050) */
void orc_potrs(int uplo, int N, int NRHS, const double *A, int lda,
               double *B, int ldb)
{
   if (!N || !NRHS) return;
   if (uplo == Upper)
   {
      orc_reftrsm(Left, Upper, Trans, NonUnit, N, NRHS, 1.0, A, lda, B, ldb);
      orc_reftrsm(Left, Upper, NoTrans, NonUnit, N, NRHS, 1.0, A, lda, B, ldb);
   }
   else
   {
      orc_reftrsm(Left, Lower, NoTrans, NonUnit, N, NRHS, 1.0, A, lda, B, ldb);
      orc_reftrsm(Left, Lower, Trans, NonUnit, N, NRHS, 1.0, A, lda, B, ldb);
   }
}

/* ------------------------------------------------------------------------
 * tester residuals
 * ------------------------------------------------------------------------ */

/* l3blastst gemmtst residual (AtlasBase/Antoine/atlbinblas.base:11791-11935,
 * gen bin/l3blastst.c:603-748):
 *   ||C - C_trusted||_1 / (max(nA,1) max(nB,1) max(nC,1) eps max(M,N,K))
 * with nA scaled by |alpha| and nC by |beta| when those exceed one, and a
 * zero norm replaced by one.  C0 is the input C. */
double orc_gemm_resid(int ta, int tb, int M, int N, int K, double alpha,
                      const double *A, int lda, const double *B, int ldb,
                      double beta, const double *C0, int ldc0,
                      const double *Ctest, int ldct, const double *Ctrust,
                      int ldtr)
{
   int mA = (ta == NoTrans) ? M : K, nA = (ta == NoTrans) ? K : M;
   int mB = (tb == NoTrans) ? K : N, nB = (tb == NoTrans) ? N : K;
   double normA, normB, normC, normD, eps = orc_epsilon();
   int mx = M > N ? M : N;
   if (K > mx) mx = K;
   if (M == 0 || N == 0) return 0.0;
   normC = orc_genrm1(M, N, C0, ldc0);
   if (fabs(beta) > 1.0) normC *= fabs(beta);
   if (normC == 0.0) normC = 1.0;
   normA = orc_genrm1(mA, nA, A, lda);
   if (fabs(alpha) > 1.0) normA *= fabs(alpha);
   if (normA == 0.0) normA = 1.0;
   normB = orc_genrm1(mB, nB, B, ldb);
   if (normB == 0.0) normB = 1.0;
   normD = orc_gediffnrm1(M, N, Ctest, ldct, Ctrust, ldtr);
   return normD / ((normA > 1.0 ? normA : 1.0) * (normB > 1.0 ? normB : 1.0) *
                   (normC > 1.0 ? normC : 1.0) * eps * mx);
}

/* gemmtst element-wise bound (AtlasBase/Clint/atlas.base:21968, gen
 * bin/gemmtst.c:395): |D-C| <= 2(|alpha| 2K eps + |beta| eps) + eps;
 * NaN in either fails.  Returns the number of bad elements. */
int orc_gemm_elem_check(int M, int N, int K, double alpha, double beta,
                        const double *Ctest, int ldct, const double *Ctrust,
                        int ldtr, double *maxerr)
{
   const double feps = DBL_EPSILON;
   const double ferr = 2.0 * (fabs(alpha) * 2.0 * K * feps + fabs(beta) * feps)
                       + feps;
   int i, j, nbad = 0;
   double mx = 0.0;
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++)
      {
         double d = Ctest[(size_t)j * ldct + i], c = Ctrust[(size_t)j * ldtr + i];
         double e = fabs(d - c);
         if (d != d || c != c || e > ferr) nbad++;
         if (e > mx) mx = e;
      }
   if (maxerr) *maxerr = mx;
   return nbad;
}

/* l3blastst trsmtst residual (atlbinblas.base:12761-12900, gen
 * bin/l3blastst.c:1525-1655):
 *   ||B - B_trusted||_1 / (max(trnrm1(A),1) max(nB,1) eps max(M,N)) */
double orc_trsm_resid(int side, int uplo, int diag, int M, int N,
                      double alpha, const double *A, int lda,
                      const double *B0, int ldb0, const double *Btest,
                      int ldbt, const double *Btrust, int ldtr)
{
   int nA = (side == Left) ? M : N;
   double normA, normB, normD, eps = orc_epsilon();
   if (M == 0 || N == 0) return 0.0;
   normB = orc_genrm1(M, N, B0, ldb0);
   if (fabs(alpha) > 1.0) normB *= fabs(alpha);
   if (normB == 0.0) normB = 1.0;
   normA = orc_trnrm1(uplo, diag, nA, A, lda);
   if (normA == 0.0) normA = 1.0;
   normD = orc_gediffnrm1(M, N, Btest, ldbt, Btrust, ldtr);
   return normD / ((normA > 1.0 ? normA : 1.0) * (normB > 1.0 ? normB : 1.0) *
                   eps * (M > N ? M : N));
}

/* lutst column-major residual (AtlasBase/Clint/atlas-tlp.base:1992-2058,
 * gen bin/lutst.c:84-252): ||P*A0 - L*U||_1 / (||A0||_1 eps min(M,N)) */
double orc_lu_resid(int M, int N, const double *A0, int lda0,
                    const double *LU, int ldlu, const int *ipiv)
{
   const int MN = M < N ? M : N;
   double *PA, *LmU, normA, resid;
   int i, j, k;
   if (M == 0 || N == 0) return 0.0;
   PA = (double *)malloc(sizeof(double) * (size_t)M * N);
   LmU = (double *)malloc(sizeof(double) * (size_t)M * N);
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++) PA[(size_t)j * M + i] = A0[(size_t)j * lda0 + i];
   normA = orc_genrm1(M, N, A0, lda0);
   orc_laswp(N, PA, M, 0, MN, ipiv, 1);
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++)
      {
         /* (L*U)(i,j) = sum_{k <= min(i,j,MN-1)} L(i,k) U(k,j), L unit */
         double s = 0.0;
         int kmax = i < j ? i : j;
         if (kmax > MN - 1) kmax = MN - 1;
         for (k = 0; k <= kmax; k++)
         {
            double l = (k == i) ? 1.0 : LU[(size_t)k * ldlu + i];
            s += l * LU[(size_t)j * ldlu + k];
         }
         LmU[(size_t)j * M + i] = s;
      }
   resid = orc_gediffnrm1(M, N, PA, M, LmU, M);
   resid /= normA * orc_epsilon() * MN;
   free(PA);
   free(LmU);
   return resid;
}

/* llttst residual (atlas-tlp.base:1450-1514, gen bin/llttst.c:33-349):
 * ||A0 - L*L^T|| (symmetric 1-norm over the stored triangle)
 *   / (lltnrm1(A0) eps N) */
double orc_llt_resid(int uplo, int N, const double *A0, int lda0,
                     const double *L, int ldl)
{
   double *C, normA, resid;
   int i, j, k;
   if (N == 0) return 0.0;
   C = (double *)calloc((size_t)N * N, sizeof(double));
   for (j = 0; j < N; j++)
   {
      if (uplo == Lower)
         for (i = j; i < N; i++)
         {
            double s = L[(size_t)j * ldl + i] * L[(size_t)j * ldl + j];
            for (k = 0; k < j; k++)
               s += L[(size_t)k * ldl + i] * L[(size_t)k * ldl + j];
            C[(size_t)j * N + i] = s - A0[(size_t)j * lda0 + i];
         }
      else
         for (i = 0; i <= j; i++)
         {
            double s = L[(size_t)i * ldl + i] * L[(size_t)j * ldl + i];
            for (k = 0; k < i; k++)
               s += L[(size_t)i * ldl + k] * L[(size_t)j * ldl + k];
            C[(size_t)j * N + i] = s - A0[(size_t)j * lda0 + i];
         }
   }
   normA = orc_lltnrm1(uplo, N, A0, lda0);
   resid = orc_lltnrm1(uplo, N, C, N) / (normA * orc_epsilon() * N);
   free(C);
   return resid;
}

/* slvtst (atlas-tlp.base:574-630, gen bin/slvtst.c:419-441):
 * max_j ||x_j - xc_j||_inf / (N eps nrmA ||xc_j||_inf) */
double orc_slv_resid(int N, int NRHS, double nrmA, const double *XC,
                     int ldxc, const double *X, int ldx)
{
   double eps = orc_epsilon(), max = 0.0;
   int i, j;
   for (j = 0; j < NRHS; j++)
   {
      double nrmX = orc_infnrm(N, XC + (size_t)j * ldxc, 1), nrmD = 0.0, t0;
      for (i = 0; i < N; i++)
      {
         double d = fabs(XC[(size_t)j * ldxc + i] - X[(size_t)j * ldx + i]);
         if (d > nrmD) nrmD = d;
      }
      t0 = nrmD / (N * eps * nrmA * nrmX);
      if (t0 > max) max = t0;
   }
   return max;
}

/* lutst NPVTS: rows that moved (gen bin/lutst.c:142-147) */
int orc_npvts(int N, const int *ipiv)
{
   int i, n = 0;
   for (i = 0; i < N; i++) if (ipiv[i] != i) n++;
   return n;
}

/* ------------------------------------------------------------------------
 * Textbook TRMM, SYMM, SYR2K (the ATL_ref* reference routines,
 * AtlasBase/Antoine/atlrefblas.base: trmm :8406+, symm :7223+, syr2k
 * :7576+) and the triangular inverse / LAUUM (LAPACK's definitions, as
 * ATL_trtri and ATL_lauum compute them, AtlasBase/Clint/atlas-lp.base:
 * 1703-2380, 221-330), by explicit sums over the referenced elements.
 * ------------------------------------------------------------------------ */

/* op(A)(i, k) of a triangular A with the given uplo / trans / diag */
static double tri_elem(int uplo, int ta, int diag, const double *A, int lda, int i, int k)
{
   const int trans = ta != NoTrans;
   const int lower_eff = (uplo == Lower) != trans;
   if (lower_eff ? i < k : i > k) return 0.0;
   if (i == k && diag == Unit) return 1.0;
   return trans ? A[k + (size_t)i * lda] : A[i + (size_t)k * lda];
}

void orc_reftrmm(int side, int uplo, int ta, int diag, int M, int N, double alpha,
                 const double *A, int lda, double *B, int ldb)
{
   double *T = (double *)malloc(sizeof(double) * (size_t)(M > 0 ? M : 1) * (N > 0 ? N : 1));
   int i, j, k;
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++) T[i + (size_t)j * M] = B[i + (size_t)j * ldb];
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++)
      {
         double t = 0.0;
         if (side == Left)
            for (k = 0; k < M; k++) t += tri_elem(uplo, ta, diag, A, lda, i, k) * T[k + (size_t)j * M];
         else
            for (k = 0; k < N; k++) t += T[i + (size_t)k * M] * tri_elem(uplo, ta, diag, A, lda, k, j);
         B[i + (size_t)j * ldb] = alpha * t;
      }
   free(T);
}

void orc_refsymm(int side, int uplo, int M, int N, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc)
{
   int i, j, k;
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++)
      {
         double t = 0.0, c;
         if (side == Left)
            for (k = 0; k < M; k++)
            {
               const int st = uplo == Lower ? i >= k : i <= k;
               const double a = st ? A[i + (size_t)k * lda] : A[k + (size_t)i * lda];
               t += a * B[k + (size_t)j * ldb];
            }
         else
            for (k = 0; k < N; k++)
            {
               const int st = uplo == Lower ? k >= j : k <= j;
               const double a = st ? A[k + (size_t)j * lda] : A[j + (size_t)k * lda];
               t += B[i + (size_t)k * ldb] * a;
            }
         c = beta == 0.0 ? 0.0 : beta * C[i + (size_t)j * ldc];
         C[i + (size_t)j * ldc] = alpha * t + c;
      }
}

void orc_refsyr2k(int uplo, int trans, int N, int K, double alpha, const double *A, int lda,
                  const double *B, int ldb, double beta, double *C, int ldc)
{
   const int tr = trans != NoTrans;
   int i, j, k;
   for (j = 0; j < N; j++)
      for (i = 0; i < N; i++)
      {
         double t = 0.0, c;
         if (uplo == Lower ? i < j : i > j) continue;
         for (k = 0; k < K; k++)
         {
            const double ai = tr ? A[k + (size_t)i * lda] : A[i + (size_t)k * lda];
            const double aj = tr ? A[k + (size_t)j * lda] : A[j + (size_t)k * lda];
            const double bi = tr ? B[k + (size_t)i * ldb] : B[i + (size_t)k * ldb];
            const double bj = tr ? B[k + (size_t)j * ldb] : B[j + (size_t)k * ldb];
            t += ai * bj + bi * aj;
         }
         c = beta == 0.0 ? 0.0 : beta * C[i + (size_t)j * ldc];
         C[i + (size_t)j * ldc] = alpha * t + c;
      }
}

/* inv(T) of the uplo triangle by substitution, column by column (T X = I);
 * a zero on a NonUnit diagonal returns its 1-based index, A unchanged. */
int orc_reftrtri(int uplo, int diag, int N, double *A, int lda)
{
   double *X;
   int i, j, k;
   if (diag == NonUnit)
      for (i = 0; i < N; i++)
         if (A[i + (size_t)i * lda] == 0.0) return i + 1;
   X = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1) * (N > 0 ? N : 1));
   for (j = 0; j < N; j++)
   {
      double *x = X + (size_t)j * N;
      for (i = 0; i < N; i++) x[i] = (i == j) ? 1.0 : 0.0;
      if (uplo == Lower)
         for (i = 0; i < N; i++)
         {
            double t = x[i];
            for (k = 0; k < i; k++) t -= A[i + (size_t)k * lda] * x[k];
            x[i] = diag == Unit ? t : t / A[i + (size_t)i * lda];
         }
      else
         for (i = N - 1; i >= 0; i--)
         {
            double t = x[i];
            for (k = i + 1; k < N; k++) t -= A[i + (size_t)k * lda] * x[k];
            x[i] = diag == Unit ? t : t / A[i + (size_t)i * lda];
         }
   }
   for (j = 0; j < N; j++)
      for (i = 0; i < N; i++)
      {
         if (uplo == Lower ? i < j : i > j) continue;
         if (i == j && diag == Unit) continue;
         A[i + (size_t)j * lda] = X[i + (size_t)j * N];
      }
   free(X);
   return 0;
}

/* Lower: A <- L^T L; Upper: A <- U U^T (the uplo triangle). */
void orc_reflauum(int uplo, int N, double *A, int lda)
{
   double *P = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1) * (N > 0 ? N : 1));
   int i, j, k;
   for (j = 0; j < N; j++)
      for (i = 0; i < N; i++)
      {
         double t = 0.0;
         if (uplo == Lower ? i < j : i > j) continue;
         if (uplo == Lower)
            for (k = i; k < N; k++) t += A[k + (size_t)i * lda] * A[k + (size_t)j * lda];
         else
            for (k = j; k < N; k++) t += A[i + (size_t)k * lda] * A[j + (size_t)k * lda];
         P[i + (size_t)j * N] = t;
      }
   for (j = 0; j < N; j++)
      for (i = 0; i < N; i++)
         if (uplo == Lower ? i >= j : i <= j) A[i + (size_t)j * lda] = P[i + (size_t)j * N];
   free(P);
}

/* ------------------------------------------------------------------------
 * The netlib CBLAS level-3 tester's machinery, restated in C (its Fortran
 * driver c_dblat3.f needs a Fortran compiler): the DBEG generator, DMAKE and
 * the DMMCH check (AtlasBase/Clint/misc/CBLATST/c_dblat3.f).
 * ------------------------------------------------------------------------ */

/* DBEG (c_dblat3.f, FUNCTION DBEG): values (i - 500)/1001 from the
 * multiplicative sequence i <- 891 i mod 1000, every fifth step skipped;
 * st[0..2] = (i, ic, mi), st[3] != 0 asks for a reset. */
double orc_blat_dbeg(int *st)
{
   if (st[3])
   {
      st[2] = 891;
      st[0] = 7;
      st[1] = 0;
      st[3] = 0;
   }
   st[1]++;
   for (;;)
   {
      st[0] = st[0] * st[2];
      st[0] = st[0] - 1000 * (st[0] / 1000);
      if (st[1] >= 5)
      {
         st[1] = 0;
         continue;
      }
      break;
   }
   return (st[0] - 500) / 1001.0;
}

/* DMAKE (c_dblat3.f): type 0 'GE', 1 'SY', 2 'TR'; fills the NMAX-leading
 * A (m x n) and the packed-by-lda copy AA, the part that must not be
 * referenced set to the rogue value -1e10. */
void orc_blat_dmake(int type, int uplo, int diag, int m, int n, double *A, int nmax,
                    double *AA, int lda, int *st, double transl)
{
   const int gen = type == 0, sym = type == 1, tri = type == 2;
   const int upper = (sym || tri) && uplo == Upper, lower = (sym || tri) && uplo == Lower;
   const int unit = tri && diag == Unit;
   int i, j;
   for (j = 0; j < n; j++)
   {
      for (i = 0; i < m; i++)
         if (gen || (upper && i <= j) || (lower && i >= j))
         {
            A[i + (size_t)j * nmax] = orc_blat_dbeg(st) + transl;
            if (i != j)
            {
               if (n > 3 && j + 1 == n / 2) A[i + (size_t)j * nmax] = 0.0;
               if (sym) A[j + (size_t)i * nmax] = A[i + (size_t)j * nmax];
               else if (tri) A[j + (size_t)i * nmax] = 0.0;
            }
         }
      if (tri) A[j + (size_t)j * nmax] += 1.0;
      if (unit) A[j + (size_t)j * nmax] = 1.0;
   }
   for (j = 0; j < n; j++)
   {
      int ibeg = 0, iend = m;  /* rows [ibeg, iend) copied, the rest rogue */
      if (!gen)
      {
         if (upper)
         {
            ibeg = 0;
            iend = unit ? j : j + 1;
         }
         else
         {
            ibeg = unit ? j + 1 : j;
            iend = n;
         }
      }
      for (i = 0; i < lda; i++)
         AA[i + (size_t)j * lda] = (i >= ibeg && i < iend) ? A[i + (size_t)j * nmax] : -1.0e10;
   }
}

/* DMMCH (c_dblat3.f): recomputes alpha op(A) op(B) + beta C column by column
 * with gauges G = |alpha| |op(A)| |op(B)| + |beta| |C| and compares with CC:
 * err_j = max_i |CT_i - CC_ij| / eps (divided by G_i when G_i != 0).  As in
 * the Fortran, *err is the last column's ratio and the result is fatal (1)
 * when some err_j * sqrt(eps) >= 1; *errmax gets the largest err_j. */
int orc_blat_dmmch(int ta, int tb, int m, int n, int kk, double alpha, const double *A, int lda,
                   const double *B, int ldb, double beta, const double *C, int ldc,
                   const double *CC, int ldcc, double eps, double *err, double *errmax)
{
   const int trana = ta != NoTrans, tranb = tb != NoTrans;
   double *ct = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
   double *g = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
   int i, j, k, fatal = 0;
   *err = 0.0;
   *errmax = 0.0;
   for (j = 0; j < n && !fatal; j++)
   {
      for (i = 0; i < m; i++) ct[i] = g[i] = 0.0;
      for (k = 0; k < kk; k++)
         for (i = 0; i < m; i++)
         {
            const double a = trana ? A[k + (size_t)i * lda] : A[i + (size_t)k * lda];
            const double b = tranb ? B[j + (size_t)k * ldb] : B[k + (size_t)j * ldb];
            ct[i] += a * b;
            g[i] += fabs(a) * fabs(b);
         }
      for (i = 0; i < m; i++)
      {
         ct[i] = alpha * ct[i] + beta * C[i + (size_t)j * ldc];
         g[i] = fabs(alpha) * g[i] + fabs(beta) * fabs(C[i + (size_t)j * ldc]);
      }
      *err = 0.0;
      for (i = 0; i < m; i++)
      {
         double erri = fabs(ct[i] - CC[i + (size_t)j * ldcc]) / eps;
         if (g[i] != 0.0) erri /= g[i];
         if (erri > *err || erri != erri) *err = erri != erri ? HUGE_VAL : erri;
         if (*err * sqrt(eps) >= 1.0)
         {
            fatal = 1;
            break;
         }
      }
      if (*err > *errmax) *errmax = *err;
   }
   free(ct);
   free(g);
   return fatal;
}

/* Wall-clock seconds of orc_refgemm (NN, alpha=beta=1) on C(:,0:ncols),
 * the columns split into nthreads contiguous slabs, one POSIX thread each
 * (each slab is the same textbook loop; the threads share nothing they
 * write). */
struct slab_arg { int M, n, K, lda, ldb, ldc; const double *A, *B; double *C; };
static void *slab_run(void *p)
{
   const struct slab_arg *a = (const struct slab_arg *)p;
   if (a->n > 0)
      orc_refgemm(NoTrans, NoTrans, a->M, a->n, a->K, 1.0, a->A, a->lda, a->B, a->ldb, 1.0,
                  a->C, a->ldc);
   return NULL;
}
double orc_time_refgemm_slab_mt(int M, int ncols, int K, const double *A, int lda,
                                const double *B, int ldb, double *C, int ldc, int nthreads)
{
   struct timespec t0, t1;
   pthread_t th[256];
   struct slab_arg arg[256];
   int t, j0 = 0;
   if (nthreads < 1) nthreads = 1;
   if (nthreads > 256) nthreads = 256;
   for (t = 0; t < nthreads; t++)
   {
      const int n = ncols / nthreads + (t < ncols % nthreads);
      arg[t].M = M; arg[t].n = n; arg[t].K = K;
      arg[t].lda = lda; arg[t].ldb = ldb; arg[t].ldc = ldc;
      arg[t].A = A; arg[t].B = B + (size_t)j0 * ldb; arg[t].C = C + (size_t)j0 * ldc;
      j0 += n;
   }
   clock_gettime(CLOCK_MONOTONIC, &t0);
   for (t = 1; t < nthreads; t++) pthread_create(&th[t], NULL, slab_run, &arg[t]);
   slab_run(&arg[0]);
   for (t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
   clock_gettime(CLOCK_MONOTONIC, &t1);
   return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}

/* Wall-clock seconds of orc_refgemm (NN, alpha=beta=1) on C(:,0:ncols) */
double orc_time_refgemm_slab(int M, int ncols, int K, const double *A,
                             int lda, const double *B, int ldb, double *C,
                             int ldc)
{
   struct timespec t0, t1;
   clock_gettime(CLOCK_MONOTONIC, &t0);
   orc_refgemm(NoTrans, NoTrans, M, ncols, K, 1.0, A, lda, B, ldb, 1.0, C, ldc);
   clock_gettime(CLOCK_MONOTONIC, &t1);
   return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
