// Placeholder for the message-driven form until it is built.
#include "jacobi2d_internal.h"
namespace chm { namespace jac {
int persistent_create(Handle*) { return fail(CHM_E_UNSUPPORTED, "persistent form not built"); }
int persistent_reset(Handle*) { return fail(CHM_E_UNSUPPORTED, "persistent form not built"); }
int persistent_solve(Handle*, chm_jacobi_result*) { return fail(CHM_E_UNSUPPORTED, "persistent form not built"); }
int persistent_read_grid(Handle*, double*) { return fail(CHM_E_UNSUPPORTED, "persistent form not built"); }
int persistent_hash_region(Handle*, int, int, int, int, uint64_t*) { return fail(CHM_E_UNSUPPORTED, "persistent form not built"); }
void persistent_destroy(Handle*) {}
}}
