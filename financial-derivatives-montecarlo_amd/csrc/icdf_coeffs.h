/* Generated by gen/fit_icdf.py -- do not edit.
 * Least-squares fit of f(w) = Phi^-1(1-q)/(1-2q), w = -log(4q(1-q)).
 * max relative fit error: central 3.126e-08, tail 1.984e-08 (FP64 evaluation). */
#pragma once
#define FDMC_ICDF_W_SPLIT 7.0f
#define FDMC_ICDF_C_CENTRAL 1.5f
#define FDMC_ICDF_C_TAIL 3.6466656632567456f
#define FDMC_ICDF_DEG_CENTRAL 11
#define FDMC_ICDF_DEG_TAIL 8
/* Horner order: highest degree first. */
#define FDMC_ICDF_CENTRAL_COEFFS {-3.068715464e-11f, 1.489208073e-09f, -2.543304672e-08f, 1.816881947e-07f, -4.649945574e-08f, -7.720461513e-06f, 3.294258050e-05f, 2.514612150e-04f, -2.959082549e-03f, 1.244069282e-03f, 3.540659990e-01f, 1.770685674e+00f}
#define FDMC_ICDF_TAIL_COEFFS {6.990816644e-05f, -2.254464726e-04f, 3.635823688e-04f, -5.374258230e-04f, 9.998705659e-04f, -2.186302909e-03f, 3.524185380e-03f, 1.425823367e+00f, 4.926125244e+00f}
/* The central polynomial in u = log2(q(1-q)): w < W_SPLIT <=> u > U_SPLIT,
 * f = sum_k d_k (u - U0)^k with d_k = c_k (-ln 2)^k. */
#define FDMC_ICDF_U0 -4.1640425613334457f
#define FDMC_ICDF_U_SPLIT -12.098865286222745f
#define FDMC_ICDF_CENTRAL_COEFFS_U {5.445486596e-13f, 3.812501230e-11f, 9.393502136e-10f, 9.681226880e-09f, 3.574588855e-09f, -8.562410177e-07f, -5.270901751e-06f, 5.804607435e-05f, 9.854474362e-04f, 5.977168361e-04f, -2.454198489e-01f, 1.770685674e+00f}
