// Matrix-free divergence template and small elementwise helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mw_solver.h"

namespace mw {

void ghat_compute(double hx, double hz, double gscale, double* Gh, hipStream_t st);
template <class T>
void gtmpl_apply_u(const BoxGeom& g, const double* Gh, const uint8_t* fix, const T* p, T* yu,
                   bool accumulate, hipStream_t st);
template <class T>
void gtmpl_apply_p(const BoxGeom& g, const double* Gh, const uint8_t* fix, const T* u, T* yp,
                   double alpha, hipStream_t st);
template <class T>
void diag_scale(long n, double a, const double* d, const T* x, T* y, hipStream_t st);

}  // namespace mw
