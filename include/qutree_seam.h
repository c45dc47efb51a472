/*
 * qutree_seam.h -- QuTree's Fortran kernel seam, served by libqutree_amd.so.
 *
 * QuTree binds the LA_lib.f kernels with these extern "C" declarations
 * (include/Core/Tensor_Implementation.h:450-459 for matvec_, ctmatvec_,
 * rmatvec_, rhomat_; include/Core/Matrix_Implementation.h:454-461 for
 * tmatvec_, trmatvec_).  The prototypes below are the same, argument for
 * argument, so that a QuTree build that drops LA_lib.f and links
 * libqutree_amd.so resolves every call site unchanged:
 *   C, B, mat: the output tensor, the input tensor and the a x a matrix,
 *              complex<double> data passed as double* (real data for
 *              rmatvec_ / trmatvec_);
 *   a, b, c:   the active, before and after dimensions, by reference;
 *   add:       Fortran logical*4, nonzero accumulates into C.
 * The pointers must be device-accessible (device or managed memory).  Each
 * call runs on the legacy default stream and returns when its result is in
 * memory, as the Fortran routine does.  The Fortran routines report nothing;
 * here a failed call leaves its output untouched, prints the reason once per
 * call to stderr and is reported by qtg_seam_status() (0 after a successful
 * call, a qtg_status otherwise) and qtg_last_error().
 */
#ifndef QUTREE_SEAM_H
#define QUTREE_SEAM_H

#ifdef __cplusplus
extern "C" {
#endif

/* mulpsi(b,k,c) (+)= matrix(k,l) psi(b,l,c)        (LA_lib.f:34-81) */
void matvec_(double *C, double *B, double *mat, int *a, int *b, int *c, int *add);
/* mulpsi(b,k,c) (+)= matrix(l,k) psi(b,l,c)        (LA_lib.f:104-151) */
void tmatvec_(double *C, double *B, double *mat, int *a, int *b, int *c, int *add);
/* mulpsi(b,k,c) (+)= conj(matrix(l,k)) psi(b,l,c)  (LA_lib.f:244-291) */
void ctmatvec_(double *C, double *B, double *mat, int *a, int *b, int *c, int *add);
/* real: mulpsi(b,k,c) (+)= matrix(k,l) psi(b,l,c)  (LA_lib.f:314-379) */
void rmatvec_(double *C, double *B, double *mat, int *a, int *b, int *c, int *add);
/* real: mulpsi(b,k,c) (+)= matrix(l,k) psi(b,l,c)  (LA_lib.f:402-467) */
void trmatvec_(double *C, double *B, double *mat, int *a, int *b, int *c, int *add);
/* matrix(j,l) = sum conj(bra(n,j,m)) ket(n,l,m); overwrites (LA_lib.f:718-762) */
void rhomat_(double *Bra, double *Ket, double *M, int *a, int *b, int *c);

/* Status of the calling thread's last seam call (0 = QTG_OK). */
int qtg_seam_status(void);

#ifdef __cplusplus
}
#endif
#endif /* QUTREE_SEAM_H */
