// Batched SCF engine entry points (see scf.hip).
#pragma once
#include "context.hpp"

namespace derek {

void scf_init(derek_ctx* c, const derek_scf_params& p);
void scf_step(derek_ctx* c, derek_scf_report& rep);
double scf_polish(derek_ctx* c, int max_iter, double res_tol);
double scf_nscf(derek_ctx* c, int max_iter, double res_tol);
// what: 0 eigenvalues (nb x nk), 1 occupations, 2 rho_out (N)
void scf_get(derek_ctx* c, int what, double* out);
// resident wavefunctions of k as complex128 npw_k x nb; put = upload
void scf_psi(derek_ctx* c, int k, double* psi, bool put);
void scf_apply_h_resident(derek_ctx* c, int k, int nb);
// out[0] k-points per batched launch, out[1] graph replay on/off
void scf_tuning(derek_ctx* c, int32_t* out);

}  // namespace derek
