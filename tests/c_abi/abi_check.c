/* A C caller of the backend's ABI: compiled with -std=c99 -Wall -Wextra -Werror -pedantic, it
 * takes every entry point of cdns.h into a table of exactly typed function pointers (so a
 * signature that drifts from the header fails to compile), links against libcdns_hip.so, and
 * checks the argument validation that needs no device. */
#include "cdns.h"

#include <stdio.h>
#include <string.h>

typedef struct {
    int (*create)(const cdns_config *, cdns_ctx **);
    void (*destroy)(cdns_ctx *);
    const char *(*last_error)(const cdns_ctx *);
    int (*sync)(cdns_ctx *);
    void *(*stream)(cdns_ctx *);
    int (*xyzfft2)(cdns_ctx *, char, double *, double *, int, int);
    int (*xyzfft)(cdns_ctx *, char, double *, double *, int, int);
    int (*to_physical)(cdns_ctx *, double *, double *, int *, int);
    int (*plane_moments)(cdns_ctx *, int, double *, int *);
    int (*diagnostics)(cdns_ctx *, double[3], int *);
    int (*curl)(cdns_ctx *, double *, const double *, int);
    int (*cross_product)(cdns_ctx *, double *, const double *, const double *, int);
    int (*nonlinear_term)(cdns_ctx *, double *, const double *, double *, int);
    int (*setup)(cdns_ctx *, double);
    int (*solve_modes)(cdns_ctx *, double *, const double *, const double *, double *, const double *,
                       const double *, int, int);
    int (*get_table)(cdns_ctx *, int, double *);
    int (*set_state)(cdns_ctx *, int, const double *, int);
    int (*get_state)(cdns_ctx *, int, double *, int);
    int (*step)(cdns_ctx *, int);
    int (*run)(cdns_ctx *, int, int);
    int (*flow_gradients)(cdns_ctx *, double[2]);
    double (*time)(const cdns_ctx *);
    int (*profile)(cdns_ctx *, int);
    int (*profile_read)(cdns_ctx *, char *, long long *, double *, int);
    int (*profile_reset)(cdns_ctx *);
    int (*set_body_force)(cdns_ctx *, const double *, int);
    int (*objective)(cdns_ctx *, int, const double *, double *, double *, int);
    int (*adjoint_step)(cdns_ctx *, int, const double *, double *, double *, double *, double *, double *, int);
    int (*adjoint_run)(cdns_ctx *, const double *, const double *, int, int, int, double *, double *, double *,
                       int);
    int (*engine_create)(unsigned long long, cdns_engine **);
    int (*engine_draw)(cdns_engine *, int, void *, unsigned long long);
    void (*engine_destroy)(cdns_engine *);
    int (*adjoint_run_noalloc)(cdns_ctx *, const double *, const double *, int, int, int, double *, double *,
                               double *, int);
} cdns_api;

static const cdns_api api = {
    cdns_create,        cdns_destroy,       cdns_last_error,     cdns_sync,          cdns_stream,
    cdns_xyzfft2,       cdns_xyzfft,        cdns_to_physical,    cdns_plane_moments, cdns_diagnostics,
    cdns_curl,          cdns_cross_product, cdns_nonlinear_term, cdns_setup,         cdns_solve_modes,
    cdns_get_table,     cdns_set_state,     cdns_get_state,      cdns_step,          cdns_run,
    cdns_flow_gradients, cdns_time,         cdns_profile,        cdns_profile_read,  cdns_profile_reset,
    cdns_set_body_force, cdns_objective,    cdns_adjoint_step,   cdns_adjoint_run,   cdns_engine_create,
    cdns_engine_draw,   cdns_engine_destroy, cdns_adjoint_run_noalloc,
};

int main(void) {
    int failures = 0;
    cdns_ctx *ctx = (cdns_ctx *)&failures; /* must be reset to NULL by a failing create */
    if (api.create(NULL, &ctx) == 0 || ctx != NULL) {
        fprintf(stderr, "cdns_create(NULL) did not fail\n");
        ++failures;
    }
    if (strcmp(api.last_error(NULL), "null context") != 0) {
        fprintf(stderr, "cdns_last_error(NULL) = %s\n", api.last_error(NULL));
        ++failures;
    }
    if (api.time(NULL) != 0.0 || api.stream(NULL) != NULL) {
        fprintf(stderr, "null-context queries\n");
        ++failures;
    }
    api.destroy(NULL);
    /* the host-only engine: std::mt19937_64's 10000th output of the default seed 5489 */
    cdns_engine *e = NULL;
    unsigned long long draws[10000];
    if (api.engine_create(5489ULL, &e) != 0 || api.engine_draw(e, 0, draws, 10000ULL) != 0 ||
        draws[9999] != 9981545732273789042ULL) {
        fprintf(stderr, "engine known answer\n");
        ++failures;
    }
    api.engine_destroy(e);
    printf("abi_check: %d entry points, %d failures\n", (int)(sizeof(api) / sizeof(void (*)(void))), failures);
    return failures != 0;
}
