// pmx_functors.hpp -- device restatements of the example coefficient callbacks
// (CoeffFct3D / BoundValueFunct3D, include/General/Constants.h:128-148) and
// of Compute_SDFEM_delta (src/FE/ConvDiff.C:520-611).  The callbacks of the
// reference are host function pointers called per cell; here they are
// selected by the PMX_PRM_EXAMPLE entry of the parameter block and evaluated
// per quadrature point on the device.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/pmx_prm.h"

namespace pmx {

struct QpCoeffs {
  double eps, b0, b1, b2, c, f;
};

__host__ __device__ inline void fourier_b(const double *prm, double x, double y, double z, double b[3]) {
  const double twopi = 6.283185307179586476925286766559;
  double dA[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  const double *F = prm + PMX_PRM_FOURIER;
  for (int c = 0; c < 3; c++)
    for (int m = 0; m < 8; m++) {
      const double *t = F + (c * 8 + m) * 5;
      const double arg = twopi * (t[1] * x + t[2] * y + t[3] * z) + t[4];
      const double cs = t[0] * cos(arg) * twopi;
      dA[c][0] += cs * t[1];
      dA[c][1] += cs * t[2];
      dA[c][2] += cs * t[3];
    }
  const double s = prm[PMX_PRM_BSCALE];
  b[0] = s * (dA[2][1] - dA[1][2]);
  b[1] = s * (dA[0][2] - dA[2][0]);
  b[2] = s * (dA[1][0] - dA[0][1]);
}

__host__ __device__ inline double grf_field(const double *prm, double x, double y, double z) {
  const double *G = prm + PMX_PRM_GRF;
  double s = 0;
  for (int k = 0; k < 64; k++) s += cos(G[4 * k] * x + G[4 * k + 1] * y + G[4 * k + 2] * z + G[4 * k + 3]);
  return s * 0.17677669529663688110021109052621;  // sqrt(2/64)
}

// CHEAP = true: only the examples whose coefficients are evaluated inline
// at every point (Laplace, sine, heat, rotating); the curl-Fourier and
// lognormal branches, which kernels for the inline examples never take, are
// left out of the code so they do not weigh on its registers
template <bool CHEAP = false>
__host__ __device__ inline QpCoeffs example_coeffs(const double *prm, double x, double y, double z) {
  const double pi = 3.14159265358979323846;
  QpCoeffs k;
  const int ex = (int)prm[PMX_PRM_EXAMPLE];
  const double eps = prm[PMX_PRM_EPS];
  k.eps = eps;
  k.c = prm[PMX_PRM_C];
  k.b0 = prm[PMX_PRM_B0];
  k.b1 = prm[PMX_PRM_B1];
  k.b2 = prm[PMX_PRM_B2];
  k.f = 0;
  const double t = prm[PMX_PRM_TIME];
  if (ex == PMX_EX_LAPLACE) {
    k.b0 = k.b1 = k.b2 = 0;
    k.c = 0;
    k.f = 3 * pi * pi * eps * sin(pi * x) * sin(pi * y) * sin(pi * z);
  } else if (ex == PMX_EX_SINE_CD) {
    const double sx = sin(pi * x), sy = sin(pi * y), sz = sin(pi * z);
    const double cx = cos(pi * x), cy = cos(pi * y), cz = cos(pi * z);
    const double u = sx * sy * sz;
    k.f = 3 * pi * pi * eps * u + pi * (k.b0 * cx * sy * sz + k.b1 * sx * cy * sz + k.b2 * sx * sy * cz) + k.c * u;
  } else if (!CHEAP && ex == PMX_EX_CURL_FOURIER) {
    if constexpr (!CHEAP) {
      double b[3];
      fourier_b(prm, x, y, z, b);
      k.b0 = b[0];
      k.b1 = b[1];
      k.b2 = b[2];
      k.c = 0;
      k.f = 1.0;
    }
  } else if (ex == PMX_EX_HEAT_AMC) {
    k.eps = 1.0;
    k.b0 = k.b1 = k.b2 = 0;
    k.c = 0;
    k.f = (3. * pi * pi - 0.1) * exp(-0.1 * t) * sin(pi * x) * cos(pi * y) * cos(pi * z);
  } else if (!CHEAP && ex == PMX_EX_LOGNORMAL) {
    if constexpr (!CHEAP) {
      k.eps = eps * exp(prm[PMX_PRM_GRF_SIGMA] * grf_field(prm, x, y, z));
      k.f = 1.0;
    }
  } else if (ex == PMX_EX_ROTATING) {
    k.b0 = -(y - 0.5);
    k.b1 = x - 0.5;
    k.b2 = 0.2;
    k.c = 0;
    k.f = 0.0;
  }
  return k;
}

__host__ __device__ inline double example_bound_value(const double *prm, double x, double y, double z) {
  const double pi = 3.14159265358979323846;
  const int ex = (int)prm[PMX_PRM_EXAMPLE];
  const double t = prm[PMX_PRM_TIME];
  if (ex == PMX_EX_LAPLACE || ex == PMX_EX_SINE_CD) return sin(pi * x) * sin(pi * y) * sin(pi * z);
  if (ex == PMX_EX_HEAT_AMC) return exp(-0.1 * t) * sin(pi * x) * cos(pi * y) * cos(pi * z);
  return 0.0;
}

__host__ __device__ inline double example_exact(const double *prm, double x, double y, double z) {
  const double pi = 3.14159265358979323846;
  const int ex = (int)prm[PMX_PRM_EXAMPLE];
  const double t = prm[PMX_PRM_TIME];
  if (ex == PMX_EX_HEAT_AMC) return exp(-0.1 * t) * sin(pi * x) * cos(pi * y) * cos(pi * z);
  if (ex == PMX_EX_ROTATING) {
    const double *B = prm + PMX_PRM_BLOB;
    const double r2 = (x - B[0]) * (x - B[0]) + (y - B[1]) * (y - B[1]) + (z - B[2]) * (z - B[2]);
    return exp(-r2 / (2 * B[3] * B[3]));
  }
  return sin(pi * x) * sin(pi * y) * sin(pi * z);
}

// the exact solution and its gradient (the ExactU functions of the examples:
// Laplace.h:12-19, amc.h:16-26; the rotating blob's reference function has a
// zero gradient entry, as in the oracle), v = {u, u_x, u_y, u_z}
__host__ __device__ inline void example_exact_grad(const double *prm, double x, double y, double z, double v[4]) {
  const double pi = 3.14159265358979323846;
  const int ex = (int)prm[PMX_PRM_EXAMPLE];
  const double t = prm[PMX_PRM_TIME];
  if (ex == PMX_EX_HEAT_AMC) {
    const double e = exp(-0.1 * t);
    v[0] = e * sin(pi * x) * cos(pi * y) * cos(pi * z);
    v[1] = pi * e * cos(pi * x) * cos(pi * y) * cos(pi * z);
    v[2] = -pi * e * sin(pi * x) * sin(pi * y) * cos(pi * z);
    v[3] = -pi * e * sin(pi * x) * cos(pi * y) * sin(pi * z);
  } else if (ex == PMX_EX_ROTATING) {
    v[0] = example_exact(prm, x, y, z);
    v[1] = v[2] = v[3] = 0.0;
  } else {
    v[0] = sin(pi * x) * sin(pi * y) * sin(pi * z);
    v[1] = pi * cos(pi * x) * sin(pi * y) * sin(pi * z);
    v[2] = pi * sin(pi * x) * cos(pi * y) * sin(pi * z);
    v[3] = pi * sin(pi * x) * sin(pi * y) * cos(pi * z);
  }
}

// Compute_SDFEM_delta for SDFEM_TYPE 0/1/2 with CELL_MEASURE 0
__host__ __device__ inline double sdfem_delta(const double *prm, double hK, double eps, double b1, double b2,
                                              double b3) {
  const double delta0 = prm[PMX_PRM_DELTA0], delta1 = prm[PMX_PRM_DELTA1];
  if (fabs(eps) < 1e-20) eps = 1e-20;
  const double nb = sqrt(b1 * b1 + b2 * b2 + b3 * b3);
  const int type = (int)prm[PMX_PRM_SDFEM_TYPE];
  double delta = 0;
  if (type == 0) {
    delta = (eps < hK * nb) ? delta0 * hK / nb : delta1 * hK * hK / eps;
  } else {
    const double pd = prm[PMX_PRM_POLYDEG] > 0 ? prm[PMX_PRM_POLYDEG] : 1.0;
    const double nu = 1.0 / pd;
    if (nb > 0) {
      const double alpha = nu * nb * hK / (2 * eps);
      delta = nu * hK * (1 / tanh(alpha) - 1 / alpha) / (2 * nb);
    }
    delta *= delta0;
  }
  return delta;
}

// The stabilisation parameter of RhsAssemble_SUPG (src/FE/TCD3D.C:396-408),
// which has its own formula: Pe = hK |b| / (2 eps), tau = hK (coth Pe - 1/Pe)
// / (2 |b|), and hK Pe / (6 |b|) for Pe < 1e-2.  With b = 0 the reference
// divides 0 by 0 (a NaN right-hand side); here the parameter is 0 there, which
// is the value every finite b.grad phi = 0 multiplies to.
__host__ __device__ inline double rhs_supg_tau(double hK, double eps, double b1, double b2, double b3) {
  const double bb = sqrt(b1 * b1 + b2 * b2 + b3 * b3);
  if (!(bb > 0)) return 0.0;
  const double pe = hK * bb / (2 * eps);
  const double tau = pe < 1e-2 ? pe / (6 * bb) : (1 / tanh(pe) - 1 / pe) / (2 * bb);
  return tau * hK;
}

// hK of a brick cell of side h for CELL_MEASURE (Assemble3D.C:314-336):
// the diameter sqrt(3) h (0 and every id without its own case), the cube root
// of the volume (1, 3) or the shortest edge (2), both h.
__host__ __device__ inline double cell_measure_hk(const double *prm, double h) {
  const int m = (int)prm[PMX_PRM_CELL_MEASURE];
  return (m == 1 || m == 2 || m == 3) ? h : sqrt(3.0) * h;
}
// Mesh_size_in_convection_direction_without_storing (ConvDiff.C:385-485) for
// an axis-parallel brick of side h: the gradients of the 8 trilinear basis
// functions at the barycentre are (sx, sy, sz) / (4h), sx, sy, sz = +-1, so
// den = sum |b . grad phi_i| and the size is 2 |b| / den (hK when den < 1e-10).
__host__ __device__ inline double hk_convection(double h, double hK, double b1, double b2, double b3) {
  double den = 0;
  for (int i = 0; i < 8; i++) {
    const double sx = (i & 1) ? 1.0 : -1.0, sy = (i & 2) ? 1.0 : -1.0, sz = (i & 4) ? 1.0 : -1.0;
    den += fabs(sx * b1 + sy * b2 + sz * b3);
  }
  den /= 4 * h;
  if (den < 1e-10) return hK;
  return 2 * sqrt(b1 * b1 + b2 * b2 + b3 * b3) / den;
}

// The quadrature-point record of the assembly, {w |detJ|, eps, b0, b1, b2, c,
// f, delta}, for a point of a brick cell of side h with rule weight w and the
// coefficients c at the point: |detJ| = h^3 / 8 (THexaAffin), hK from
// CELL_MEASURE, delta the streamline parameter of the form (0 for the forms
// without one).  c0: the coefficients at the cell's first quadrature point;
// with CELL_MEASURE 4 the reference computes the mesh size in the convection
// direction at the cell's first call of Compute_SDFEM_delta (its point 0)
// and keeps it for the cell (INTERNAL_HK_CONVECTION, ConvDiff.C:373-383,
// reset per cell in Assemble3D.C:389-405).
__host__ __device__ inline void qp_record(const double *prm, int form, double h, double w, const QpCoeffs &c,
                                          double *o, const QpCoeffs *c0 = nullptr) {
  const double hK = cell_measure_hk(prm, h);
  const bool conv = (int)prm[PMX_PRM_CELL_MEASURE] == 4;
  const QpCoeffs &f = c0 ? *c0 : c;
  double delta = 0;
  if (form == PMX_FORM_SDFEM || form == PMX_FORM_MASS_SUPG) {
    // BilinearAssemble_SD and MatrixMRhsAssemble_SUPG (TCD3D.C:103-104):
    // Compute_SDFEM_delta of the unscaled coefficients
    const double hd = conv ? hk_convection(h, hK, f.b0, f.b1, f.b2) : hK;
    delta = sdfem_delta(prm, hd, c.eps, c.b0, c.b1, c.b2);
  } else if (form == PMX_FORM_RHS_SUPG) {
    // its own parameter with the cell measure (no Compute_SDFEM_delta)
    delta = rhs_supg_tau(hK, c.eps, c.b0, c.b1, c.b2);
  } else if (form == PMX_FORM_TCD_SUPG_AK) {
    // MatricesAKRhsAssemble_SUPG (TCD3D.C:167-181, 206-208): delta of the
    // theta1 tau scaled coefficients, times theta1 tau
    const double v = prm[PMX_PRM_THETA1] * prm[PMX_PRM_TAU];
    const double hd = conv ? hk_convection(h, hK, v * f.b0, v * f.b1, v * f.b2) : hK;
    delta = sdfem_delta(prm, hd, v * c.eps, v * c.b0, v * c.b1, v * c.b2) * v;
  }
  o[0] = w * (h * h * h / 8.0);
  o[1] = c.eps; o[2] = c.b0; o[3] = c.b1; o[4] = c.b2; o[5] = c.c; o[6] = c.f; o[7] = delta;
}
}  // namespace pmx
