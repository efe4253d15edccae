/*
 * atl_oracle.h -- CPU oracle for the FP64 level-3 path (TEST INFRASTRUCTURE).
 *
 * This is a plain-C restatement of ATLAS 3.11.42's test generator, norms,
 * tester residual formulas and textbook (reference) routines.  It is only
 * ever used as the checker: by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg.  No product code links or calls it.
 *
 * Every function cites the reference basefile:line it restates (paths are
 * relative to the ATLAS source tree).  Enum values are ATLAS's own
 * (AtlasBase/Clint/atlas.base:19795+): NoTrans=111 Trans=112 ConjTrans=113,
 * Upper=121 Lower=122, NonUnit=131 Unit=132, Left=141 Right=142.
 */
#ifndef ATL_ORACLE_H
#define ATL_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* --- deterministic generator: AtlasBase/Clint/atlas-tst.base:5353-5620 --- */
void   orc_srand(int iseed);
int    orc_rand(void);
double orc_dumb_rand(void);                       /* atlas_tst.h dumb_rand */
void   orc_rand_words(int iseed, int n, int *out); /* n successive ATL_rand */
void   orc_gegen(int M, int N, double *A, int lda, int seed);
int    orc_gechkgap(int M, int N, const double *A, int lda);
void   orc_trddom(int uplo, int N, double *A, int lda);
void   orc_lltgen(int uplo, int N, double *A, int lda, int seed);

/* --- scalars and norms --- */
double orc_epsilon(void);
double orc_genrm1(int M, int N, const double *A, int lda);
double orc_gediffnrm1(int M, int N, const double *A, int lda,
                      const double *B, int ldb);
double orc_trnrm1(int uplo, int diag, int N, const double *A, int lda);
double orc_lltnrm1(int uplo, int N, const double *A, int lda);
double orc_infnrm(int N, const double *x, int incx);

/* --- textbook routines (column-major) --- */
void orc_refgemm(int ta, int tb, int M, int N, int K, double alpha,
                 const double *A, int lda, const double *B, int ldb,
                 double beta, double *C, int ldc);
void orc_refsyrk(int uplo, int trans, int N, int K, double alpha,
                 const double *A, int lda, double beta, double *C, int ldc);
void orc_reftrsm(int side, int uplo, int ta, int diag, int M, int N,
                 double alpha, const double *A, int lda, double *B, int ldb);
int  orc_idamax(int N, const double *x, int incx);
int  orc_getf2(int M, int N, double *A, int lda, int *ipiv);
void orc_laswp(int N, double *A, int lda, int K1, int K2, const int *ipiv,
               int inci);
int  orc_potf2(int uplo, int N, double *A, int lda);
void orc_getrs(int trans, int N, int NRHS, const double *A, int lda,
               const int *ipiv, double *B, int ldb);
void orc_potrs(int uplo, int N, int NRHS, const double *A, int lda,
               double *B, int ldb);

/* --- tester residuals --- */
double orc_gemm_resid(int ta, int tb, int M, int N, int K, double alpha,
                      const double *A, int lda, const double *B, int ldb,
                      double beta, const double *C0, int ldc0,
                      const double *Ctest, int ldct, const double *Ctrust,
                      int ldtr);
int    orc_gemm_elem_check(int M, int N, int K, double alpha, double beta,
                           const double *Ctest, int ldct,
                           const double *Ctrust, int ldtr, double *maxerr);
double orc_trsm_resid(int side, int uplo, int diag, int M, int N,
                      double alpha, const double *A, int lda,
                      const double *B0, int ldb0, const double *Btest,
                      int ldbt, const double *Btrust, int ldtr);
double orc_lu_resid(int M, int N, const double *A0, int lda0,
                    const double *LU, int ldlu, const int *ipiv);
double orc_llt_resid(int uplo, int N, const double *A0, int lda0,
                     const double *L, int ldl);
double orc_slv_resid(int N, int NRHS, double nrmA, const double *XC,
                     int ldxc, const double *X, int ldx);
int    orc_npvts(int N, const int *ipiv);

/* --- CPU baseline helper: wall time of refgemm on a column slab --- */
void orc_reftrmm(int side, int uplo, int ta, int diag, int M, int N, double alpha,
                 const double *A, int lda, double *B, int ldb);
void orc_refsymm(int side, int uplo, int M, int N, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc);
void orc_refsyr2k(int uplo, int trans, int N, int K, double alpha, const double *A, int lda,
                  const double *B, int ldb, double beta, double *C, int ldc);
int orc_reftrtri(int uplo, int diag, int N, double *A, int lda);
void orc_reflauum(int uplo, int N, double *A, int lda);
double orc_blat_dbeg(int *st);
void orc_blat_dmake(int type, int uplo, int diag, int m, int n, double *A, int nmax,
                    double *AA, int lda, int *st, double transl);
int orc_blat_dmmch(int ta, int tb, int m, int n, int kk, double alpha, const double *A, int lda,
                   const double *B, int ldb, double beta, const double *C, int ldc,
                   const double *CC, int ldcc, double eps, double *err, double *errmax);
double orc_time_refgemm_slab_mt(int M, int ncols, int K, const double *A, int lda,
                                const double *B, int ldb, double *C, int ldc, int nthreads);
double orc_time_refgemm_slab(int M, int ncols, int K, const double *A,
                             int lda, const double *B, int ldb, double *C,
                             int ldc);

#ifdef __cplusplus
}
#endif
#endif
