/* A C caller of libpardiso2_amd.so through include/pardiso2_amd.h, with the
 * reference's own prototypes of the routines the library replaces
 * (parallel/head3mg.c:350-377) declared beside the header: the compiler
 * rejects the file if the two disagree.  Exercises the host-side entry
 * points without a device: matclear/gausmat, the row store of gausmat2 on
 * bound arrays, the context-bound routines' error codes, default options. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pardiso2_amd.h"

/* the reference's declarations (parallel/head3mg.c:350-377) */
int gausmat2(int, long int, long int, long int *, double *);
void gausmat(int, double *, double *);
void matclear(int, double *);
void vpiterate(int, int);
void ronurestrict(int);
void vpprolong(int, int);
void vprestrict(int);
int gausmat4(int, long int, long int, int);

static int fails = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);          \
      fails++;                                                        \
    }                                                                 \
  } while (0)

int main(void) {
  /* matclear + gausmat on a 3 x 4 augmented system, solution (1, -2, 3) */
  double a[12], x[3];
  for (int i = 0; i < 12; i++) a[i] = 99.0;
  matclear(3, a);
  for (int i = 0; i < 12; i++) CHECK(a[i] == 0.0);
  const double A[3][3] = {{4, 1, 0}, {1, 5, 2}, {0, 2, 6}}, s[3] = {1, -2, 3};
  for (int i = 0; i < 3; i++) {
    double b = 0;
    for (int j = 0; j < 3; j++) {
      a[i * 4 + j] = A[i][j];
      b += A[i][j] * s[j];
    }
    a[i * 4 + 3] = b;
  }
  gausmat(3, a, x);
  for (int i = 0; i < 3; i++) CHECK(fabs(x[i] - s[i]) < 1e-13);

  /* gausmat2 needs the caller's arrays */
  long un1[8];
  double ui1[8];
  un1[0] = 3; un1[1] = 5; un1[2] = 9; un1[3] = 1;
  ui1[0] = 7.5; ui1[1] = 2.0; ui1[2] = -1.0; ui1[3] = 0.25;
  CHECK(gausmat2(1, 5, 10, un1, ui1) == STK_ENOCTX);

  enum { NR = 16, NV = 64 };
  double *sol0 = calloc(NR, sizeof(double)), *fre0 = calloc(NR, sizeof(double)), *fre1 = calloc(NR, sizeof(double));
  double *val1 = calloc(NV, sizeof(double));
  int *lin0 = calloc(NV, sizeof(int)), *num0 = calloc(NR, sizeof(int)), *pos0 = calloc(NR, sizeof(int));
  stk_ref_arrays ra;
  memset(&ra, 0, sizeof ra);
  ra.sol0 = sol0; ra.fre0 = fre0; ra.fre1 = fre1; ra.val1 = val1;
  ra.lin0 = lin0; ra.num0 = num0;
  CHECK(stk_bind_arrays(&ra) == STK_EINVAL); /* pos0 missing */
  ra.pos0 = pos0;
  CHECK(stk_bind_arrays(&ra) == STK_OK);

  /* two rows as the reference's row builders hand them over (un1[0] entries,
     ui1[0] the right-hand side), stored from the given positions */
  CHECK(gausmat2(1, 5, 10, un1, ui1) == 0);
  un1[0] = 2; un1[1] = 6; un1[2] = 2;
  ui1[0] = -1.0; ui1[1] = 1.0; ui1[2] = -0.5;
  CHECK(gausmat2(1, 6, 13, un1, ui1) == 0);
  CHECK(fre1[5] == 7.5 && pos0[5] == 10 && num0[5] == 3);
  CHECK(lin0[10] == 5 && lin0[11] == 9 && lin0[12] == 1);
  CHECK(val1[10] == 2.0 && val1[11] == -1.0 && val1[12] == 0.25);
  CHECK(fre1[6] == -1.0 && pos0[6] == 13 && num0[6] == 2);
  CHECK(lin0[13] == 6 && lin0[14] == 2 && val1[13] == 1.0 && val1[14] == -0.5);
  /* a later call for a row replaces it */
  un1[0] = 1; un1[1] = 5; ui1[0] = 3.0; ui1[1] = 4.0;
  CHECK(gausmat2(1, 5, 20, un1, ui1) == 0);
  CHECK(fre1[5] == 3.0 && pos0[5] == 20 && num0[5] == 1 && lin0[20] == 5 && val1[20] == 4.0);

  /* the device routines need a context selected by stk_bind */
  CHECK(stk_bind(NULL) == STK_OK);
  CHECK(gausmat4(1, 0, 0, 0) == STK_ENOCTX);
  vprestrict(1);
  vpprolong(1, 0);
  ronurestrict(1);
  vpiterate(0, 0);

  stk_solver_opts o;
  stk_default_opts(&o, 3);
  CHECK(o.precision == 0 && o.restart == 128 && o.tol == 1e-8 && o.krylov == 1 && o.schur == 1);
  CHECK(o.pre[0] == 3 && o.post[0] == 3 && o.ksweeps == 3 && o.basis == 1);

  CHECK(stk_bind_arrays(NULL) == STK_OK);
  CHECK(gausmat2(1, 5, 10, un1, ui1) == STK_ENOCTX);

  free(sol0); free(fre0); free(fre1); free(val1); free(lin0); free(num0); free(pos0);
  if (fails) {
    printf("%d checks failed\n", fails);
    return 1;
  }
  printf("ok\n");
  return 0;
}
