/* An unmodified BLAS caller: one dgemm_ C = 2*A*B + C on integer data (exact in FP64), checked
 * against its own loop. Linked against libtmm_blas.so, alone or with a host BLAS after it.
 * Prints "BLAS CALLER OK" and exits 0 when C is right. Test infrastructure only. */
#include <stdio.h>
#include <stdlib.h>

void dgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k, const double* alpha,
            const double* a, const int* lda, const double* b, const int* ldb, const double* beta, double* c,
            const int* ldc);

int main(void) {
  const int m = 300, n = 200, k = 250;
  double* a = malloc(sizeof(double) * m * k);
  double* b = malloc(sizeof(double) * k * n);
  double* c = malloc(sizeof(double) * m * n);
  double* r = malloc(sizeof(double) * m * n);
  for (int i = 0; i < m * k; ++i) a[i] = (i * 7) % 10;
  for (int i = 0; i < k * n; ++i) b[i] = (i * 3) % 10;
  for (int i = 0; i < m * n; ++i) c[i] = r[i] = i % 10;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double s = 0;
      for (int p = 0; p < k; ++p) s += a[i + p * m] * b[p + j * k];
      r[i + j * m] = 2 * s + r[i + j * m];
    }
  const double alpha = 2, beta = 1;
  dgemm_("N", "N", &m, &n, &k, &alpha, a, &m, b, &k, &beta, c, &m);
  int ok = 1;
  for (int i = 0; i < m * n && ok; ++i) ok = c[i] == r[i];
  printf(ok ? "BLAS CALLER OK\n" : "BLAS CALLER WRONG\n");
  return ok ? 0 : 1;
}
