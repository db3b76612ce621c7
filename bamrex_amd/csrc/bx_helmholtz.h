// Internal interface between the Helmholtz solver and the learned smoother.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/bamrex_hip.h"

// e = omega * (sigma / D) * N(r / sigma) on an nx x ny periodic field, with
// sigma = rms(r) and D the operator diagonal registered by bx_cnn_set_operator.
int bx_cnn_forward(bx_cnn *net, const double *r, double *e, int nx, int ny,
                   double omega, hipStream_t stream);

// Allocates the inference buffers for npx pixels now (a graph capture that
// runs bx_cnn_forward must not allocate).
int bx_cnn_reserve(bx_cnn *net, long long npx);
// Changes whenever a buffer bx_cnn_forward uses is reallocated.
long long bx_cnn_epoch(const bx_cnn *net);

// Records msg as the text bx_helm_last_error returns and gives back code
// (shared by the bx_helm_* and bx_imex_* entry points).
int bx_helm_fail(int code, const std::string &msg);
