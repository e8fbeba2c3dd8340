// les_duct_kernels.hip - duct-mode scheme (periodic duct, SSP-RK3, table
// properties, level-set G).  Not built yet: every entry reports LES_ERR_PARAM.
#include "les_engine.hpp"

struct DuctState {
    int unused;
};

int  duct_create(les_gpu* c) { (void)c; return LES_ERR_PARAM; }
void duct_destroy(les_gpu* c) { (void)c; }
int  duct_init_state(les_gpu* c) { (void)c; return LES_ERR_PARAM; }
int  duct_upload_state(les_gpu* c, const double* const* f, int nf) { (void)c; (void)f; (void)nf; return LES_ERR_PARAM; }
int  duct_download_state(les_gpu* c, double* const* f, int nf) { (void)c; (void)f; (void)nf; return LES_ERR_PARAM; }
int  duct_download_field(les_gpu* c, const char* name, double* out) { (void)c; (void)name; (void)out; return LES_ERR_PARAM; }
int  duct_step(les_gpu* c, int n_steps, les_step_report* rep) { (void)c; (void)n_steps; (void)rep; return LES_ERR_PARAM; }
int  duct_rhs(les_gpu* c, double* const* out, int n_out) { (void)c; (void)out; (void)n_out; return LES_ERR_PARAM; }
void duct_params_changed(les_gpu* c) { (void)c; }
