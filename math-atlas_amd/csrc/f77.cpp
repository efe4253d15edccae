// Fortran 77 interface of the level-3 hot path: the DGEMM, DSYRK, DTRSM,
// DTRMM, DSYMM, DSYR2K BLAS and the DGETRF, DGETRS, DGESV, DPOTRF, DPOTRS,
// DPOSV, DLASWP, DTRTRI, DPOTRI, DLAUUM LAPACK routines with the gfortran
// calling convention (lower case, trailing underscore, every argument by
// reference; the hidden CHARACTER lengths a Fortran caller appends are not
// read).  Pivots are 1-based, as in Fortran.
//
// Replaces ATLAS's F77 layer: the Fortran argument checks of the BLAS
// wrappers (AtlasBase/Antoine/atlF77blas.base: DGEMM :1585-1679, DSYRK
// :1861-1953, DTRSM/DTRMM :1954-2060, DSYMM :1680-1764, DSYR2K :1765-1860;
// XERBLA :115-140, LSAME :55-114), the C bodies they call
// (AtlasBase/Antoine/atlF77wrap.base, e.g. ATL_F77wrap_gemm :1026-1090
// with its A==B SYRK route) and the LAPACK wrappers
// (AtlasBase/Clint/atlas-fint.base: POSV :1044, POTRF :1169, POTRS :1289,
// GETRS :1395, GESV :1518, GETRF :1616, TRTRI :1703, POTRI :1825,
// LAUUM :2038).  The checks are restated in C: the same order, the same
// INFO values, XERBLA with the routine's name.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "atlas_gpu_atl.h"
#include "atlas_gpu_clapack.h"
#include "atlas_gpu_f77.h"
#include "atlg_internal.hpp"

namespace {

// LSAME: case-insensitive comparison of one character (ASCII).
bool lsame(const char* ca, char cb) {
  char a = *ca;
  if (a >= 'a' && a <= 'z') a = static_cast<char>(a - 32);
  if (cb >= 'a' && cb <= 'z') cb = static_cast<char>(cb - 32);
  return a == cb;
}

int imax1(int a) { return a > 1 ? a : 1; }

void xerbla(const char* name, int info) {
  xerbla_(name, &info, 6);
}

int trans_code(const char* t) {
  if (lsame(t, 'N')) return CblasNoTrans;
  if (lsame(t, 'T')) return CblasTrans;
  if (lsame(t, 'C')) return CblasConjTrans;
  return 0;
}


// The DTRSM / DTRMM checks (atlF77blas.base:1954-2060); 0 when legal.
int trxm_f77_checks(const char* side, const char* uplo, const char* transa, const char* diag,
                    const int* m, const int* n, const int* lda, const int* ldb) {
  int info = 0, nrowa = 0;
  if (lsame(side, 'L')) nrowa = *m;
  else if (lsame(side, 'R')) nrowa = *n;
  else info = 1;
  if (!lsame(uplo, 'U') && !lsame(uplo, 'L') && info == 0) info = 2;
  if (!trans_code(transa) && info == 0) info = 3;
  if (!lsame(diag, 'N') && !lsame(diag, 'U') && info == 0) info = 4;
  if (info == 0) {
    if (*m < 0) info = 5;
    else if (*n < 0) info = 6;
    else if (*lda < imax1(nrowa)) info = 9;
    else if (*ldb < imax1(*m)) info = 11;
  }
  return info;
}

std::vector<int> zero_based(const int* ipiv, int n) {
  std::vector<int> p(n > 0 ? n : 1);
  for (int i = 0; i < n; i++) p[i] = ipiv[i] - 1;
  return p;
}

// The UPLO / N / LDA checks of DPOTRF, DPOTRI and DLAUUM (INFO -1, -2, -4).
int potrf_f77_checks(const char* uplo, const int* n, const int* lda) {
  if (!lsame(uplo, 'U') && !lsame(uplo, 'L')) return -1;
  if (*n < 0) return -2;
  if (*lda < imax1(*n)) return -4;
  return 0;
}
CBLAS_UPLO uplo_of(const char* u) { return lsame(u, 'U') ? CblasUpper : CblasLower; }

// The DPOTRS / DPOSV checks (INFO -1, -2, -3, -5, -7).
int potrs_f77_checks(const char* uplo, const int* n, const int* nrhs, const int* lda,
                     const int* ldb) {
  if (!lsame(uplo, 'U') && !lsame(uplo, 'L')) return -1;
  if (*n < 0) return -2;
  if (*nrhs < 0) return -3;
  if (*lda < imax1(*n)) return -5;
  if (*ldb < imax1(*n)) return -7;
  return 0;
}

}  // namespace

extern "C" {

// XERBLA (atlF77blas.base:115-140): prints the reference message and
// stops.  Weak, so a program may supply its own; a hook installed with
// atlg_set_xerbla_hook receives the call instead (p = INFO).
__attribute__((weak)) void xerbla_(const char* srname, const int* info, size_t len) {
  char name[8] = {0};
  for (size_t i = 0; i < 6 && i < (len ? len : 6) && srname[i]; i++) name[i] = srname[i];
  char msg[96];
  snprintf(msg, sizeof msg, " ** On entry to %-6s parameter number %2d had an illegal value\n",
           name, *info);
  atlg::f77_xerbla(*info, name, msg);
}

void dgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k,
            const double* alpha, const double* a, const int* lda, const double* b,
            const int* ldb, const double* beta, double* c, const int* ldc) {
  int info = 0, nrowa = 0, nrowb = 0;
  const int ta = trans_code(transa), tb = trans_code(transb);
  if (ta == CblasNoTrans) nrowa = *m;
  else if (ta) nrowa = *k;
  else info = 1;
  if (tb == CblasNoTrans) nrowb = *k;
  else if (tb) nrowb = *n;
  else if (info == 0) info = 2;
  if (info == 0) {
    if (*m < 0) info = 3;
    else if (*n < 0) info = 4;
    else if (*k < 0) info = 5;
    else if (*lda < imax1(nrowa)) info = 8;
    else if (*ldb < imax1(nrowb)) info = 10;
    else if (*ldc < imax1(*m)) info = 13;
  }
  if (info) {
    xerbla("DGEMM ", info);
    return;
  }
  // cblas_dgemm routes A == B, M == N, TA != TB, lda == ldb, beta == 0 to
  // SYRK plus a reflection, as ATL_F77wrap_gemm does
  cblas_dgemm(CblasColMajor, (CBLAS_TRANSPOSE)ta, (CBLAS_TRANSPOSE)tb, *m, *n, *k, *alpha, a,
              *lda, b, *ldb, *beta, c, *ldc);
}

void dsyrk_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha,
            const double* a, const int* lda, const double* beta, double* c, const int* ldc) {
  int info = 0, nrowa = 0;
  const bool up = lsame(uplo, 'U');
  if (!up && !lsame(uplo, 'L')) info = 1;
  const int t = trans_code(trans);
  if (t == CblasNoTrans) nrowa = *n;
  else if (t) nrowa = *k;
  else if (info == 0) info = 2;
  if (info == 0) {
    if (*n < 0) info = 3;
    else if (*k < 0) info = 4;
    else if (*lda < imax1(nrowa)) info = 7;
    else if (*ldc < imax1(*n)) info = 10;
  }
  if (info) {
    xerbla("DSYRK ", info);
    return;
  }
  cblas_dsyrk(CblasColMajor, up ? CblasUpper : CblasLower, (CBLAS_TRANSPOSE)t, *n, *k, *alpha, a,
              *lda, *beta, c, *ldc);
}


void dtrsm_(const char* side, const char* uplo, const char* transa, const char* diag,
            const int* m, const int* n, const double* alpha, const double* a, const int* lda,
            double* b, const int* ldb) {
  const int info = trxm_f77_checks(side, uplo, transa, diag, m, n, lda, ldb);
  if (info) {
    xerbla("DTRSM ", info);
    return;
  }
  cblas_dtrsm(CblasColMajor, lsame(side, 'L') ? CblasLeft : CblasRight,
              lsame(uplo, 'U') ? CblasUpper : CblasLower, (CBLAS_TRANSPOSE)trans_code(transa),
              lsame(diag, 'U') ? CblasUnit : CblasNonUnit, *m, *n, *alpha, a, *lda, b, *ldb);
}

void dtrmm_(const char* side, const char* uplo, const char* transa, const char* diag,
            const int* m, const int* n, const double* alpha, const double* a, const int* lda,
            double* b, const int* ldb) {
  const int info = trxm_f77_checks(side, uplo, transa, diag, m, n, lda, ldb);
  if (info) {
    xerbla("DTRMM ", info);
    return;
  }
  cblas_dtrmm(CblasColMajor, lsame(side, 'L') ? CblasLeft : CblasRight,
              lsame(uplo, 'U') ? CblasUpper : CblasLower, (CBLAS_TRANSPOSE)trans_code(transa),
              lsame(diag, 'U') ? CblasUnit : CblasNonUnit, *m, *n, *alpha, a, *lda, b, *ldb);
}

// DSYMM (atlF77blas.base:1680-1764)
void dsymm_(const char* side, const char* uplo, const int* m, const int* n, const double* alpha,
            const double* a, const int* lda, const double* b, const int* ldb,
            const double* beta, double* c, const int* ldc) {
  int info = 0, nrowa = 0;
  if (lsame(side, 'L')) nrowa = *m;
  else if (lsame(side, 'R')) nrowa = *n;
  else info = 1;
  if (!lsame(uplo, 'U') && !lsame(uplo, 'L') && info == 0) info = 2;
  if (info == 0) {
    if (*m < 0) info = 3;
    else if (*n < 0) info = 4;
    else if (*lda < imax1(nrowa)) info = 7;
    else if (*ldb < imax1(*m)) info = 9;
    else if (*ldc < imax1(*m)) info = 12;
  }
  if (info) {
    xerbla("DSYMM ", info);
    return;
  }
  cblas_dsymm(CblasColMajor, lsame(side, 'L') ? CblasLeft : CblasRight,
              lsame(uplo, 'U') ? CblasUpper : CblasLower, *m, *n, *alpha, a, *lda, b, *ldb,
              *beta, c, *ldc);
}

// DSYR2K (atlF77blas.base:1765-1860)
void dsyr2k_(const char* uplo, const char* trans, const int* n, const int* k,
             const double* alpha, const double* a, const int* lda, const double* b,
             const int* ldb, const double* beta, double* c, const int* ldc) {
  int info = 0, nrowa = 0;
  if (!lsame(uplo, 'U') && !lsame(uplo, 'L')) info = 1;
  const int t = trans_code(trans);
  if (t == CblasNoTrans) nrowa = *n;
  else if (t) nrowa = *k;
  else if (info == 0) info = 2;
  if (info == 0) {
    if (*n < 0) info = 3;
    else if (*k < 0) info = 4;
    else if (*lda < imax1(nrowa)) info = 7;
    else if (*ldb < imax1(nrowa)) info = 9;
    else if (*ldc < imax1(*n)) info = 12;
  }
  if (info) {
    xerbla("DSYR2K", info);
    return;
  }
  cblas_dsyr2k(CblasColMajor, lsame(uplo, 'U') ? CblasUpper : CblasLower, (CBLAS_TRANSPOSE)t,
               *n, *k, *alpha, a, *lda, b, *ldb, *beta, c, *ldc);
}

// DGETRF (atlas-fint.base:1616-1702): ATL_dgetrf, then the pivots 1-based.
void dgetrf_(const int* m, const int* n, double* a, const int* lda, int* ipiv, int* info) {
  *info = 0;
  if (*m < 0) *info = -1;
  else if (*n < 0) *info = -2;
  else if (*lda < imax1(*m)) *info = -4;
  if (*info) {
    xerbla("DGETRF", -*info);
    return;
  }
  *info = ATL_dgetrf(CblasColMajor, *m, *n, a, *lda, ipiv);
  const int mn = *m < *n ? *m : *n;
  for (int i = 0; i < mn; i++) ipiv[i] += 1;
}


// DGETRS (atlas-fint.base:1395-1517)
void dgetrs_(const char* trans, const int* n, const int* nrhs, const double* a, const int* lda,
             const int* ipiv, double* b, const int* ldb, int* info) {
  *info = 0;
  const int t = trans_code(trans);
  if (!t) *info = -1;
  else if (*n < 0) *info = -2;
  else if (*nrhs < 0) *info = -3;
  else if (*lda < imax1(*n)) *info = -5;
  else if (*ldb < imax1(*n)) *info = -8;
  if (*info) {
    xerbla("DGETRS", -*info);
    return;
  }
  const std::vector<int> p = zero_based(ipiv, *n);
  ATL_dgetrs(CblasColMajor, (CBLAS_TRANSPOSE)t, *n, *nrhs, a, *lda, p.data(), b, *ldb);
}

// DGESV (atlas-fint.base:1518-1615)
void dgesv_(const int* n, const int* nrhs, double* a, const int* lda, int* ipiv, double* b,
            const int* ldb, int* info) {
  *info = 0;
  if (*n < 0) *info = -1;
  else if (*nrhs < 0) *info = -2;
  else if (*lda < imax1(*n)) *info = -4;
  else if (*ldb < imax1(*n)) *info = -7;
  if (*info) {
    xerbla("DGESV ", -*info);
    return;
  }
  *info = ATL_dgetrf(CblasColMajor, *n, *n, a, *lda, ipiv);
  if (*info == 0) ATL_dgetrs(CblasColMajor, CblasNoTrans, *n, *nrhs, a, *lda, ipiv, b, *ldb);
  for (int i = 0; i < *n; i++) ipiv[i] += 1;
}


// DPOTRF (atlas-fint.base:1169-1288)
void dpotrf_(const char* uplo, const int* n, double* a, const int* lda, int* info) {
  *info = potrf_f77_checks(uplo, n, lda);
  if (*info) {
    xerbla("DPOTRF", -*info);
    return;
  }
  *info = ATL_dpotrf(CblasColMajor, uplo_of(uplo), *n, a, *lda);
}


// DPOTRS (atlas-fint.base:1289-1394)
void dpotrs_(const char* uplo, const int* n, const int* nrhs, const double* a, const int* lda,
             double* b, const int* ldb, int* info) {
  *info = potrs_f77_checks(uplo, n, nrhs, lda, ldb);
  if (*info) {
    xerbla("DPOTRS", -*info);
    return;
  }
  ATL_dpotrs(CblasColMajor, uplo_of(uplo), *n, *nrhs, a, *lda, b, *ldb);
}

// DPOSV (atlas-fint.base:1044-1168)
void dposv_(const char* uplo, const int* n, const int* nrhs, double* a, const int* lda,
            double* b, const int* ldb, int* info) {
  *info = potrs_f77_checks(uplo, n, nrhs, lda, ldb);
  if (*info) {
    xerbla("DPOSV ", -*info);
    return;
  }
  *info = ATL_dpotrf(CblasColMajor, uplo_of(uplo), *n, a, *lda);
  if (*info == 0) ATL_dpotrs(CblasColMajor, uplo_of(uplo), *n, *nrhs, a, *lda, b, *ldb);
}

// DTRTRI (atlas-fint.base:1703-1824)
void dtrtri_(const char* uplo, const char* diag, const int* n, double* a, const int* lda,
             int* info) {
  *info = 0;
  if (!lsame(uplo, 'U') && !lsame(uplo, 'L')) *info = -1;
  else if (!lsame(diag, 'N') && !lsame(diag, 'U')) *info = -2;
  else if (*n < 0) *info = -3;
  else if (*lda < imax1(*n)) *info = -5;
  if (*info) {
    xerbla("DTRTRI", -*info);
    return;
  }
  *info = ATL_dtrtri(CblasColMajor, uplo_of(uplo), lsame(diag, 'U') ? CblasUnit : CblasNonUnit,
                     *n, a, *lda);
}

// DPOTRI (atlas-fint.base:1825-1937)
void dpotri_(const char* uplo, const int* n, double* a, const int* lda, int* info) {
  *info = potrf_f77_checks(uplo, n, lda);
  if (*info) {
    xerbla("DPOTRI", -*info);
    return;
  }
  *info = ATL_dpotri(CblasColMajor, uplo_of(uplo), *n, a, *lda);
}

// DLAUUM (atlas-fint.base:2038-2142)
void dlauum_(const char* uplo, const int* n, double* a, const int* lda, int* info) {
  *info = potrf_f77_checks(uplo, n, lda);
  if (*info) {
    xerbla("DLAUUM", -*info);
    return;
  }
  ATL_dlauum(CblasColMajor, uplo_of(uplo), *n, a, *lda);
}

// DLASWP (LAPACK semantics, no argument checks): rows K1..K2 (1-based)
// exchanged with IPIV(K1 + (I-K1)*|INCX|) in increasing order for INCX > 0,
// decreasing for INCX < 0; nothing for INCX = 0.
void dlaswp_(const int* n, double* a, const int* lda, const int* k1, const int* k2,
             const int* ipiv, const int* incx) {
  if (*incx == 0 || *k2 < *k1) return;
  const int inc = *incx > 0 ? *incx : -*incx;
  std::vector<int> p(*k2);
  for (int i = *k1; i <= *k2; i++) p[i - 1] = ipiv[*k1 - 1 + (i - *k1) * inc] - 1;
  ATL_dlaswp(*n, a, *lda, *k1 - 1, *k2, p.data(), *incx > 0 ? 1 : -1);
}

}  // extern "C"
