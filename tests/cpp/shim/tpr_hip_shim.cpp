// src/tpr_hip_shim.cpp -- replaces src/tpr.cu
#include <cstdio>
#include <cstdlib>
#include "tpr.cuh"
#include "tpr_hip.h"

namespace {
void die(tprh_status st) {  // the reference exits on any device error (src/tpr.cu)
    std::fprintf(stderr, "[Error] %s (%s)\n", tprh_last_error(), tprh_status_string(st));
    std::exit(EXIT_FAILURE);
}
}  // namespace

void TPR_CU::tpr_cu(float *a, float *c, float *rhs, float *x, int n, int s) {
    tprh_status st = tprh_tpr_cu(a, c, rhs, x, n, s);
    if (st != TPRH_SUCCESS) die(st);
}
void TPR_CU::cr_cu(float *a, float *c, float *rhs, float *x, int n) {
    tprh_status st = tprh_cr_cu(a, c, rhs, x, n);
    if (st != TPRH_SUCCESS) die(st);
}
