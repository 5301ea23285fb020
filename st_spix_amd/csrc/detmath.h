/*
 * Deterministic double-precision natural logarithm.
 *
 * The split/merge Hastings ratios and the log-determinant of the shape
 * covariance decide discrete outcomes (accept/reject, argmax of the boundary
 * update).  The device math library and the host libm may differ in the last
 * bit, so both the device kernels and the CPU checker evaluate log() through
 * this one function, which uses only IEEE-754 +, -, *, / (no fused
 * multiply-add; every translation unit that includes it is compiled with
 * -ffp-contract=off) and bit manipulation.  The result is therefore the same
 * bit pattern on the GPU and on the host.
 *
 * Method: x = m * 2^e with m in [sqrt(1/2), sqrt(2)); with s = (m-1)/(m+1),
 * log(m) = 2*atanh(s) = 2*(s + s^3/3 + s^5/5 + ...)  (Abramowitz & Stegun
 * 4.1.27 / 4.5.22), |s| <= 0.1716, truncated after s^27 (relative term size
 * < 1e-21); log(x) = e*ln2_hi + (log(m) + e*ln2_lo) with ln2 split so that
 * e*ln2_hi is exact for |e| < 2^11.  Accuracy is within a few ulp of the
 * correctly rounded value; tests pin it against libm.
 */
#ifndef STSP_DETMATH_H
#define STSP_DETMATH_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define STSP_HD __host__ __device__
#else
#define STSP_HD
#endif

static inline STSP_HD double stsp_bits_to_f64(uint64_t b) {
  double d;
  memcpy(&d, &b, sizeof(d));
  return d;
}
static inline STSP_HD uint64_t stsp_f64_to_bits(double d) {
  uint64_t b;
  memcpy(&b, &d, sizeof(b));
  return b;
}

static inline STSP_HD double stsp_det_log(double x) {
  if (x != x) return x;                                    /* NaN */
  if (x < 0.0) return stsp_bits_to_f64(0x7ff8000000000000ull);
  if (x == 0.0) return stsp_bits_to_f64(0xfff0000000000000ull); /* -inf */
  uint64_t b = stsp_f64_to_bits(x);
  int ex = (int)((b >> 52) & 0x7ff);
  if (ex == 0x7ff) return x;                               /* +inf */
  int e = 0;
  if (ex == 0) {                                           /* subnormal */
    x = x * 18014398509481984.0;                           /* 2^54 */
    b = stsp_f64_to_bits(x);
    ex = (int)((b >> 52) & 0x7ff);
    e = -54;
  }
  e += ex - 1022;
  /* m in [0.5, 1) */
  double m = stsp_bits_to_f64((b & 0x800fffffffffffffull) | (0x3feull << 52));
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e -= 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  /* 1/3 + s2/5 + s2^2/7 + ... + s2^12/27 */
  double p = 1.0 / 27.0;
  p = 1.0 / 25.0 + s2 * p;
  p = 1.0 / 23.0 + s2 * p;
  p = 1.0 / 21.0 + s2 * p;
  p = 1.0 / 19.0 + s2 * p;
  p = 1.0 / 17.0 + s2 * p;
  p = 1.0 / 15.0 + s2 * p;
  p = 1.0 / 13.0 + s2 * p;
  p = 1.0 / 11.0 + s2 * p;
  p = 1.0 / 9.0 + s2 * p;
  p = 1.0 / 7.0 + s2 * p;
  p = 1.0 / 5.0 + s2 * p;
  p = 1.0 / 3.0 + s2 * p;
  double logm = 2.0 * s + (2.0 * s) * (s2 * p);
  const double ln2_hi = 6.93147180369123816490e-01; /* 0x3FE62E42FEE00000 */
  const double ln2_lo = 1.90821492927058770002e-10; /* 0x3DEA39EF35793C76 */
  double de = (double)e;
  return de * ln2_hi + (logm + de * ln2_lo);
}

/*
 * Shape (inverse-Wishart) log marginal of one part of a label, the term of
 * marginal_likelihood_shape_p (split_merge_prop.cu:1596-1621):
 *   nu * log det Psi - (nu + n) * log det Sigma_hat
 * with Sigma_hat the sample covariance (divided by n) of the part's n pixel
 * positions from their exact integer sums, Psi the prior scatter
 * (xx, xy, yy) and nu the prior count; -1e7 when det Sigma_hat < 0.001, as
 * the reference.  The lgamma terms the reference computes but does not use
 * are left out.  Shared by the device kernels and the CPU checker so both
 * give the same bits.
 */
static inline STSP_HD double stsp_shape_marginal(int64_t n, const int64_t s[5], const double psi[3],
                                                 double nu) {
  const double dn = (double)n;
  const double mx = (double)s[0] / dn, my = (double)s[1] / dn;
  const double cxx = (double)s[2] / dn - mx * mx;
  const double cxy = (double)s[3] / dn - mx * my;
  const double cyy = (double)s[4] / dn - my * my;
  double det_est = cxx * cyy - cxy * cxy;
  det_est = det_est < 0.0 ? -det_est : det_est;
  if (det_est < 0.001) return -10000000.0;
  double det_pr = psi[0] * psi[2] - psi[1] * psi[1];
  det_pr = det_pr < 0.0 ? -det_pr : det_pr;
  return nu * stsp_det_log(det_pr) - (nu + dn) * stsp_det_log(det_est);
}

/* The split's shape term: parts k and s of label k against the whole (the
 * sums of both parts); 0 when the prior scatter is degenerate
 * (det Psi < 0.001, where the reference asserts). */
static inline STSP_HD double stsp_split_shape_term(int64_t nk, const int64_t sk[5], int64_t ns,
                                                   const int64_t ss[5], const double psi[3],
                                                   double nu) {
  double det_pr = psi[0] * psi[2] - psi[1] * psi[1];
  det_pr = det_pr < 0.0 ? -det_pr : det_pr;
  if (det_pr < 0.001) return 0.0;
  int64_t sf[5];
  for (int c = 0; c < 5; c++) sf[c] = sk[c] + ss[c];
  return stsp_shape_marginal(nk, sk, psi, nu) + stsp_shape_marginal(ns, ss, psi, nu) -
         stsp_shape_marginal(nk + ns, sf, psi, nu);
}

#endif /* STSP_DETMATH_H */
