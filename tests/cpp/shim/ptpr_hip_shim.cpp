// src/ptpr_hip_shim.cpp -- replaces src/ptpr.cu
#include <cstdio>
#include <cstdlib>
#include "ptpr.cuh"
#include "tpr_hip.h"

namespace {
void die(tprh_status st) {  // the reference exits on any device error (src/ptpr.cu:509-517)
    std::fprintf(stderr, "[Error] %s (%s)\n", tprh_last_error(), tprh_status_string(st));
    std::exit(EXIT_FAILURE);
}
}  // namespace

void PTPR_CU::ptpr_cu(float *a, float *c, float *rhs, float *x, int n, int s) {
    tprh_status st = tprh_ptpr_cu(a, c, rhs, x, n, s);
    if (st != TPRH_SUCCESS) die(st);
}
void PTPR_CU::pcr_cu(float *a, float *c, float *rhs, float *x, int n) {
    tprh_status st = tprh_pcr_cu(a, c, rhs, x, n);
    if (st != TPRH_SUCCESS) die(st);
}
