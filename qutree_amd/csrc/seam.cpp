// QuTree's Fortran kernel seam (include/qutree_seam.h) on the device kernels:
// each routine forwards its by-reference dimensions to the qtg_* node entry
// point of the same meaning and waits for the result, as the Fortran routine
// returns with its output written.
#include <cstdio>

#include "../../include/qutree_seam.h"
#include "common.hpp"

namespace {
thread_local int g_seam_status = QTG_OK;

void finish(const char* name, int rc) {
  if (rc == QTG_OK) rc = (hipStreamSynchronize(nullptr) == hipSuccess) ? QTG_OK : QTG_EDEVICE;
  g_seam_status = rc;
  if (rc != QTG_OK) std::fprintf(stderr, "qutree_amd %s: %s\n", name, qtg_last_error());
}
}  // namespace

extern "C" {

void matvec_(double* C, double* B, double* mat, int* a, int* b, int* c, int* add) {
  finish("matvec_", qtg_matvec_z(C, B, mat, *a, *b, *c, *add, nullptr));
}
void tmatvec_(double* C, double* B, double* mat, int* a, int* b, int* c, int* add) {
  finish("tmatvec_", qtg_tmatvec_z(C, B, mat, *a, *b, *c, *add, nullptr));
}
void ctmatvec_(double* C, double* B, double* mat, int* a, int* b, int* c, int* add) {
  finish("ctmatvec_", qtg_ctmatvec_z(C, B, mat, *a, *b, *c, *add, nullptr));
}
void rmatvec_(double* C, double* B, double* mat, int* a, int* b, int* c, int* add) {
  finish("rmatvec_", qtg_rmatvec_d(C, B, mat, *a, *b, *c, *add, nullptr));
}
void trmatvec_(double* C, double* B, double* mat, int* a, int* b, int* c, int* add) {
  finish("trmatvec_", qtg_trmatvec_d(C, B, mat, *a, *b, *c, *add, nullptr));
}
void rhomat_(double* Bra, double* Ket, double* M, int* a, int* b, int* c) {
  finish("rhomat_", qtg_rhomat_z(Bra, Ket, M, *a, *b, *c, nullptr));
}

int qtg_seam_status(void) { return g_seam_status; }

}  // extern "C"
