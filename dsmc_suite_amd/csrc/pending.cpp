/*
 * pending.cpp - entry points of include/dsmc_hip.h whose device path is not
 * built yet.  Each returns DSMC_ERR_ARG with a message; none computes anything
 * on the host.
 */
#include <stdio.h>
#include "../../include/dsmc_hip.h"

static int pending(const char *what)
{
    fprintf(stderr, "dsmc_hip: %s is not available in this build\n", what);
    return DSMC_ERR_ARG;
}

extern "C" {
int dsmc_run_batch(int, const dsmc_params *, const dsmc_surface *, dsmc_result *, const dsmc_launch_shape *)
{ return pending("dsmc_run_batch"); }
int dsmc_autotune(const dsmc_params *, const dsmc_surface *, dsmc_launch_shape *)
{ return pending("dsmc_autotune"); }
dsmc_server *dsmc_server_create(size_t, int) { pending("dsmc_server_create"); return NULL; }
int dsmc_server_submit(dsmc_server *, const dsmc_params *, const dsmc_surface *, dsmc_result *, uint64_t *)
{ return pending("dsmc_server_submit"); }
int dsmc_server_poll(dsmc_server *, uint64_t) { return pending("dsmc_server_poll"); }
int dsmc_server_wait(dsmc_server *, uint64_t) { return pending("dsmc_server_wait"); }
int dsmc_server_cancel(dsmc_server *, uint64_t) { return pending("dsmc_server_cancel"); }
int dsmc_server_step(dsmc_server *) { return pending("dsmc_server_step"); }
void dsmc_server_destroy(dsmc_server *) {}
}
