/* Generated by gen/fit_icdf.py -- do not edit.
 * Least-squares fit of f(w) = Phi^-1(1-q)/(1-2q), w = -log(4q(1-q)).
 * max relative fit error: central 1.904e-08, tail 3.985e-08 (FP64 evaluation). */
#pragma once
#define FDMC_ICDF_W_SPLIT 5.0f
#define FDMC_ICDF_C_CENTRAL 2.5f
#define FDMC_ICDF_C_TAIL 3.4418239964743451f
#define FDMC_ICDF_DEG_CENTRAL 9
#define FDMC_ICDF_DEG_TAIL 10
/* Horner order: highest degree first. */
#define FDMC_ICDF_CENTRAL_COEFFS {-1.280012780e-08f, 3.240326213e-08f, 6.639885504e-07f, -4.891965033e-06f, -7.039139210e-06f, 3.087684964e-04f, -1.771614185e-03f, -5.907698848e-03f, 3.488020256e-01f, 2.123313476e+00f}
#define FDMC_ICDF_TAIL_COEFFS {-3.204192021e-05f, 5.138257253e-05f, 1.131192011e-04f, -4.248310562e-04f, 7.488828857e-04f, -1.167977390e-03f, 1.859675706e-03f, -3.319562956e-03f, 5.176711183e-03f, 1.424064983e+00f, 4.634225870e+00f}
