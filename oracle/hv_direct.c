/*
 * hv_direct.c -- CPU restatement of spMatVec_direct_main
 * (src/DMRG_SUPERBLOCK_SETUP.f90:740-865).  TEST INFRASTRUCTURE ONLY: the
 * checker and the timed CPU baseline; the product never links it.
 *
 * Every operator block is a row-list sparse matrix as in the reference's
 * sparse_matrix type (src/MATRIX/MATRIX_SPARSE.f90:29-71): for row r the
 * column indices cols[rowptr[r]..rowptr[r+1]) and values vals[...].  The
 * loops and their order are the reference's:
 *   1_L x H_R  (:767-779), H_L x 1_R (:784-796), then for each term
 *   C = B.V (:819-832) into a scratch allocated per term per call (:812),
 *   Hv += C.A^T (:839-853).
 * Superblock vector index: i = ir + il*Dr + off  (0-based form of :770).
 */
#include <stdlib.h>
#include <string.h>

typedef struct {
  int nrow, ncol;
  const int* rowptr; /* nrow+1 */
  const int* cols;
  const double* vals;
} csr_t;

void hv_direct_csr(int nq, const int* Dl, const int* Dr, const long* off, const csr_t* HL, const csr_t* HR,
                   int nterms, const int* dst, const int* src, const csr_t* A, const csr_t* B,
                   const double* v, double* hv, long n) {
  memset(hv, 0, sizeof(double) * (size_t)n);
  for (int k = 0; k < nq; ++k) {
    const int dl = Dl[k], dr = Dr[k];
    const long o = off[k];
    const csr_t* hr = &HR[k];
    for (int il = 0; il < dl; ++il)
      for (int ir = 0; ir < dr; ++ir) {
        const long i = ir + (long)il * dr + o;
        double acc = hv[i];
        for (int p = hr->rowptr[ir]; p < hr->rowptr[ir + 1]; ++p)
          acc += hr->vals[p] * v[hr->cols[p] + (long)il * dr + o];
        hv[i] = acc;
      }
    const csr_t* hl = &HL[k];
    for (int ir = 0; ir < dr; ++ir)
      for (int il = 0; il < dl; ++il) {
        const long i = ir + (long)il * dr + o;
        double acc = hv[i];
        for (int p = hl->rowptr[il]; p < hl->rowptr[il + 1]; ++p)
          acc += hl->vals[p] * v[ir + (long)hl->cols[p] * dr + o];
        hv[i] = acc;
      }
  }
  for (int t = 0; t < nterms; ++t) {
    const csr_t* a = &A[t];
    const csr_t* b = &B[t];
    const long rowoff = off[dst[t]], coloff = off[src[t]];
    double* C = (double*)calloc((size_t)b->nrow * (size_t)a->ncol, sizeof(double));
    for (int aj = 0; aj < a->ncol; ++aj)
      for (int bi = 0; bi < b->nrow; ++bi) {
        double acc = 0.0;
        for (int p = b->rowptr[bi]; p < b->rowptr[bi + 1]; ++p)
          acc += b->vals[p] * v[b->cols[p] + (long)aj * b->ncol + coloff];
        C[bi + (size_t)aj * b->nrow] += acc;
      }
    for (int bi = 0; bi < b->nrow; ++bi)
      for (int ai = 0; ai < a->nrow; ++ai) {
        const long i = bi + (long)ai * b->nrow + rowoff;
        double acc = hv[i];
        for (int p = a->rowptr[ai]; p < a->rowptr[ai + 1]; ++p)
          acc += a->vals[p] * C[bi + (size_t)a->cols[p] * b->nrow];
        hv[i] = acc;
      }
    free(C);
  }
}
