/* A plain column-major dgemm_ (ops 'N' and 'T'/'C', BLAS argument order) that stands in for "the
 * next BLAS" behind libtmm_blas.so in the interposer's failure tests. Test infrastructure only.
 * It announces each call on stdout so that a test can tell which library ran the product. */
#include <stdio.h>

void dgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k, const double* alpha,
            const double* a, const int* lda, const double* b, const int* ldb, const double* beta, double* c,
            const int* ldc) {
  const int tra = *ta == 'T' || *ta == 't' || *ta == 'C' || *ta == 'c';
  const int trb = *tb == 'T' || *tb == 't' || *tb == 'C' || *tb == 'c';
  printf("naive dgemm_ called\n");
  fflush(stdout);
  for (int j = 0; j < *n; ++j)
    for (int i = 0; i < *m; ++i) {
      double s = 0;
      for (int p = 0; p < *k; ++p) {
        const double av = tra ? a[p + (long)i * *lda] : a[i + (long)p * *lda];
        const double bv = trb ? b[j + (long)p * *ldb] : b[p + (long)j * *ldb];
        s += av * bv;
      }
      double* cp = &c[i + (long)j * *ldc];
      *cp = *alpha == 0 ? (*beta == 0 ? 0 : *beta * *cp) : (*beta == 0 ? *alpha * s : *alpha * s + *beta * *cp);
    }
}
