/*
 * pmx_prm.h -- flat parameter block shared by the C-ABI (libpmx) and its callers.
 *
 * ParMooN keeps every assembly/solver parameter in the global struct
 * TDatabase::ParamDB / TimeDB (include/General/Database.h:55-1066) and the
 * example callbacks read them implicitly.  The C-ABI takes them explicitly as
 * one `double prm[PMX_PRM_LEN]` block so that calls are reentrant and the
 * same block can be handed to every entry point.
 *
 * Index map (all values are doubles; integer-valued entries are rounded):
 */
#ifndef PMX_PRM_H
#define PMX_PRM_H

enum pmx_prm_index {
  PMX_PRM_EXAMPLE = 0,     /* coefficient functor, enum pmx_example            */
  PMX_PRM_EPS = 1,         /* diffusion eps = 1/RE_NR (Examples/CD_3D/Laplace.h:37) */
  PMX_PRM_B0 = 2,          /* constant convection b (examples with constant b) */
  PMX_PRM_B1 = 3,
  PMX_PRM_B2 = 4,
  PMX_PRM_C = 5,           /* reaction c                                        */
  PMX_PRM_TIME = 6,        /* TimeDB->CURRENTTIME                               */
  PMX_PRM_FORM = 7,        /* local form, enum pmx_form                         */
  PMX_PRM_SDFEM_TYPE = 8,  /* ParamDB->SDFEM_TYPE (src/General/Database.C:480)  */
  PMX_PRM_DELTA0 = 9,      /* ParamDB->DELTA0                                   */
  PMX_PRM_DELTA1 = 10,     /* ParamDB->DELTA1                                   */
  PMX_PRM_TAU = 11,        /* TimeDB->CURRENTTIMESTEPLENGTH                     */
  PMX_PRM_THETA1 = 12,     /* TimeDB->THETA1                                    */
  PMX_PRM_LINEAR = 13,     /* ParamDB->INTERNAL_PROBLEM_LINEAR (quad rule 2*deg instead of 3*deg) */
  PMX_PRM_BSCALE = 14,     /* normalisation of the curl-of-potential convection */
  PMX_PRM_POLYDEG = 15,    /* ParamDB->INTERNAL_POLYNOMIAL_DEGREE (Database.C:914) */
  PMX_PRM_FOURIER = 16,    /* 3 components x 8 modes x {amp, kx, ky, kz, phase} = 120 */
  PMX_PRM_GRF = 136,       /* 64 features x {wx, wy, wz, phase} = 256           */
  PMX_PRM_GRF_SIGMA = 392, /* log-normal field: eps(x) = eps * exp(sigma*G(x))  */
  PMX_PRM_BLOB = 393,      /* Gaussian blob centre (3) and width (1)            */
  PMX_PRM_CELL_MEASURE = 397, /* ParamDB->CELL_MEASURE (Assemble3D.C:314-336): hK of the
                               SUPG parameters; 0 diameter, 1 / 3 cube root of the volume,
                               2 shortest edge, 4 mesh size in the convection direction
                               (ConvDiff.C:373-485)                              */
  PMX_PRM_LEN = 512
};

/* Coefficient functors.  Each one restates an example header's CoeffFct3D,
 * BoundValueFunct3D and exact solution (Examples/CD_3D, Examples/TCD_3D). */
enum pmx_example {
  PMX_EX_LAPLACE = 0,      /* Examples/CD_3D/Laplace.h: u = sin sin sin, b = 0   */
  PMX_EX_SINE_CD = 1,      /* u = sin sin sin with constant b, f = -eps Lap u + b.grad u + c u */
  PMX_EX_CURL_FOURIER = 2, /* b = bscale * curl(sum of seeded Fourier modes), f = 1, g = 0 */
  PMX_EX_HEAT_AMC = 3,     /* Examples/TCD_3D/amc.h: u = e^{-0.1t} sin(pi x) cos(pi y) cos(pi z) */
  PMX_EX_LOGNORMAL = 4,    /* eps(x) = eps*exp(sigma G(x)), constant b, f = 1, g = 0 */
  PMX_EX_ROTATING = 5      /* b = (-(y-1/2), x-1/2, 0.2), f = 0, g = 0, blob initial value */
};

/* Local assembling forms (the AssembleFct3D callbacks). */
enum pmx_form {
  PMX_FORM_GALERKIN = 1,    /* BilinearAssemble      src/FE/ConvDiff3D.C:87-146  */
  PMX_FORM_SDFEM = 2,       /* BilinearAssemble_SD   src/FE/ConvDiff3D.C:150-214 */
  PMX_FORM_MASS = 3,        /* MatrixMRhsAssemble    src/FE/TCD3D.C:37-72        */
  PMX_FORM_TCD_A = 4,       /* MatrixARhsAssemble    src/FE/TCD3D.C:268-322      */
  PMX_FORM_TCD_SUPG_AK = 5, /* MatricesAKRhsAssemble_SUPG src/FE/TCD3D.C:128-266 (A and K) */
  PMX_FORM_MASS_SUPG = 6,   /* MatrixMRhsAssemble_SUPG src/FE/TCD3D.C:74-126: Galerkin mass,
                               rhs f (phi + delta b.grad phi)                    */
  PMX_FORM_RHS_SUPG = 7,    /* RhsAssemble_SUPG      src/FE/TCD3D.C:374-421 (rhs only) */
  PMX_FORM_RHS = 8          /* RhsAssemble           src/FE/TCD3D.C:423-446 (rhs only) */
};

#endif
