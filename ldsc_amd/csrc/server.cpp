// server.cpp -- request server (placeholder until the serving loop lands).
#include "internal.h"

extern "C" int ldsc_server_create(const ldsc_server_config*, ldsc_server**, char* err, uint64_t n) {
    return ldsc::fail(err, n, LDSC_ERR_INVALID_ARG, "server not available yet");
}
extern "C" int ldsc_server_destroy(ldsc_server*) { return LDSC_OK; }
extern "C" int ldsc_l2_submit(ldsc_server*, const ldsc_panel*, const ldsc_snp*, uint64_t,
                              const ldsc_l2_config*, const double*, uint64_t, double*, double*,
                              uint64_t*, char* err, uint64_t n) {
    return ldsc::fail(err, n, LDSC_ERR_INVALID_ARG, "server not available yet");
}
extern "C" int ldsc_l2_poll(ldsc_server*, uint64_t, double*, char* err, uint64_t n) {
    return ldsc::fail(err, n, LDSC_ERR_NOT_FOUND, "server not available yet");
}
extern "C" int ldsc_l2_wait(ldsc_server*, uint64_t, double*, char* err, uint64_t n) {
    return ldsc::fail(err, n, LDSC_ERR_NOT_FOUND, "server not available yet");
}
