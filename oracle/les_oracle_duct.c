/*
 * les_oracle_duct.c - CPU oracle of the duct-mode scheme (test infrastructure).
 * Placeholder until the duct scheme lands: every entry reports LES_ERR_PARAM.
 */
#include "les_oracle_internal.h"

int  ora_duct_create(ora_ctx* c) { (void)c; return LES_ERR_PARAM; }
void ora_duct_destroy(ora_ctx* c) { (void)c; }
int  ora_duct_init_state(ora_ctx* c) { (void)c; return LES_ERR_PARAM; }
int  ora_duct_step(ora_ctx* c, les_step_report* rep) { (void)c; (void)rep; return LES_ERR_PARAM; }
int  ora_duct_rhs(ora_ctx* c, double* const* out, int n_out) { (void)c; (void)out; (void)n_out; return LES_ERR_PARAM; }
int  ora_duct_set_state(ora_ctx* c, const double* const* f, int nf) { (void)c; (void)f; (void)nf; return LES_ERR_PARAM; }
int  ora_duct_get_state(const ora_ctx* c, double* const* f, int nf) { (void)c; (void)f; (void)nf; return LES_ERR_PARAM; }
int  ora_duct_get_field(const ora_ctx* c, const char* name, double* out) { (void)c; (void)name; (void)out; return LES_ERR_PARAM; }
