// cp_server.hip -- serving loop (continuous batching of independent problems).
#include <hip/hip_runtime.h>

#include "cp_problem.h"

extern "C" {

int cp_server_create(cp_ctx*, const cp_server_config*, cp_server** out)
{
    if (out) *out = nullptr;
    return cp::fail(CP_ERR_STATE, "serving loop not built yet");
}
int cp_server_destroy(cp_server*) { return CP_OK; }
int cp_server_submit(cp_server*, int, int, const double*, const double*, const int*, const double*, const double*,
                     const unsigned char*, const cp_material*, int, double, const cp_options*, long long*)
{
    return cp::fail(CP_ERR_STATE, "serving loop not built yet");
}
int cp_server_step(cp_server*, int*) { return cp::fail(CP_ERR_STATE, "serving loop not built yet"); }
int cp_server_result(cp_server*, long long, cp_request_info*, double*, double*)
{
    return cp::fail(CP_ERR_STATE, "serving loop not built yet");
}
int cp_server_clock(cp_server*, double*) { return cp::fail(CP_ERR_STATE, "serving loop not built yet"); }

}  // extern "C"
