/*
 * Scalar physics shared by the device kernels and the host exports:
 * the two boundary-flux models of the reference plugin table, the RyR
 * propensities, and the Philox4x32-10 generator used for gating.
 */
#ifndef CS_PHYSICS_H
#define CS_PHYSICS_H

#include <stdint.h>
#include <hip/hip_runtime.h>

#define CS_HD __host__ __device__ inline

/* K_0 of the exact two-voxel relaxation in ryr_flux (boundaryfluxes.c:
 * 127-130): the unitary current i_c is normalised to the hard-coded
 * 1300-0.1 uM gradient, z = 2, F = 96485 C/mol. */
CS_HD double cs_ryr_k0(double dt, double h, double i_c)
{
  return exp(-2 * dt * 1e15 / (2 * 96485.) * i_c / (1300. - 0.1) / (h * h * h));
}

/* ryr_flux: J = (u0-u1)(1-K_0)/2, the amount that moves from the SR
 * voxel u0 to the cleft voxel u1 in one step. */
CS_HD double cs_ryr_flux_d(double dt, double h, double u0, double u1, double i_c)
{
  double K_0 = cs_ryr_k0(dt, h, i_c);
  return (u0 - u1) * (1 - K_0) / 2;
}

/* serca_flux (boundaryfluxes.c:12-85): three-state SERCA cycle rate with
 * the reference's rate constants, converted from pumps/s to uM/ms over a
 * voxel face of area h^2 into a voxel of volume h^3.  u0 is cytosolic and
 * u1 SR calcium (uM); density in pumps per nm^2. */
CS_HD double cs_serca_flux_d(double dt, double h, double u0, double u1, double density)
{
  const double Ca_sr = u1 / 1000, Ca_i = u0 / 1000;            /* mM */
  const double k_p1 = 25900, k_p2 = 2540, k_p3 = 20.5;
  const double k_m1 = 16, k_m2 = 67200, k_m3 = 149;
  const double kdcai = 0.9, kdcasr = 2.24;
  const double kdh1 = 1.09e-5, kdhi = 3.54e-3, kdhsr = 1.05e-8, kdh = 7.24e-5;
  const double H_i = 1e-4, ATP = 5, ADP = 36.3e-3, P_i = 1;
  const double T_Cai = Ca_i / kdcai, T_Casr = Ca_sr / kdcasr;
  const double T_Cai2 = T_Cai * T_Cai, T_Casr2 = T_Casr * T_Casr;
  const double T_H1 = H_i / kdh1, T_Hi = (H_i * H_i) / kdhi;
  const double T_Hsr = (H_i * H_i) / kdhsr, T_H = H_i / kdh;
  const double den_i = T_Cai2 + T_Cai2 * T_Hi + T_Hi * (1 + T_H1);
  const double den_sr = T_Casr2 * T_H + T_H + T_Hsr * (1 + T_H);
  const double a_p1 = k_p1 * ATP;
  const double a_p2 = k_p2 * T_Cai2 / den_i;
  const double a_p3 = k_p3 * T_Hsr / den_sr;
  const double a_m1 = k_m1 * T_Hi / den_i;
  const double a_m2 = k_m2 * ADP * T_Casr2 * T_H / den_sr;
  const double a_m3 = k_m3 * P_i;
  const double s1 = a_p2 * a_p3 + a_m1 * a_p3 + a_m1 * a_m2;
  const double s2 = a_p1 * a_p3 + a_m2 * a_p1 + a_m2 * a_m3;
  const double s3 = a_p1 * a_p2 + a_m3 * a_m1 + a_m3 * a_p2;
  const double v_cycle = (a_p1 * a_p2 * a_p3 - a_m1 * a_m2 * a_m3) / (s1 + s2 + s3);
  const double per_pump = v_cycle * (1e-3 / 6e17);   /* umol/ms per pump   */
  const double mass = (h * h) * (density * per_pump); /* umol/ms over a face*/
  const double vol = 1e-24 * h * h * h;               /* litres             */
  return dt * (mass / vol);
}

/* Clamped power-law rates of evaluate_RyR_stochastically
 * (boundaryfluxes.c:339-342). p uses the reference parameter order. */
CS_HD double cs_rate_close(double c, const double* p)
{
  return fmax(fmin(pow(c / p[0], p[5]), p[1]), p[7]);
}
CS_HD double cs_rate_open(double c, const double* p)
{
  return fmax(fmin(pow(c / p[9], p[2]), p[3]), p[4]);
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 * as easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies,
 * Weyl key schedule. */
CS_HD void cs_philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

/* Gating uniform in [0,1) with 53 random bits: key (seed, instance),
 * counter (time index low/high word, gating group, channel). */
CS_HD double cs_gate_uniform(uint32_t seed, uint32_t instance, uint64_t ti,
                             uint32_t group, uint32_t channel)
{
  uint32_t c[4] = {(uint32_t)ti, (uint32_t)(ti >> 32), group, channel};
  cs_philox(c, seed, instance);
  const uint64_t m = ((uint64_t)c[0] << 21) | (c[1] >> 11);
  return (double)m * 0x1p-53;
}

#endif /* CS_PHYSICS_H */
