// Colour preprocessing of the command-line path (demo.cu:248-250):
// uint8 -> float rescale and the sRGB <-> CIE Lab conversions of
// rgb2lab.cu:21-181, with the reference's channel conventions and mixed
// precision: arithmetic in double, the power functions in float, Lab scaled
// to L/-100, a/100, b/100.  The forward conversion reads B,G,R (OpenCV
// order) and the inverse writes R,G,B.  A pixel with a NaN channel is left
// unwritten, as in the reference.
#include "kernels.h"

namespace stsp {

__global__ void k_rescale_u8(const uint8_t* __restrict__ in, float* __restrict__ out, long long n,
                             float scale) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = (float)(in[i] / scale);  // int / float division in float (rgb2lab.cu:172-174)
}

__device__ __forceinline__ double srgb_to_linear(double v) {
  return v <= 0.04045 ? v / 12.92 : (double)powf((float)((v + 0.055) / 1.055), 2.4f);
}

__device__ __forceinline__ double lab_f(double t) {
  const double epsilon = 0.008856, kappa = 903.3;
  return t > epsilon ? (double)powf((float)t, (float)(1.0 / 3.0)) : (kappa * t + 16.0) / 116.0;
}

__global__ void k_rgb_to_lab(const float* __restrict__ bgr, float* __restrict__ lab, long long n) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double sB = bgr[3 * t], sG = bgr[3 * t + 1], sR = bgr[3 * t + 2];
  if (sR != sR || sG != sG || sB != sB) return;
  const double r = srgb_to_linear(sR), g = srgb_to_linear(sG), b = srgb_to_linear(sB);
  const double X = r * 0.4124564 + g * 0.3575761 + b * 0.1804375;
  const double Y = r * 0.2126729 + g * 0.7151522 + b * 0.0721750;
  const double Z = r * 0.0193339 + g * 0.1191920 + b * 0.9503041;
  const double fx = lab_f(X / 0.950456), fy = lab_f(Y / 1.0), fz = lab_f(Z / 1.088754);
  const float lval = (float)(116.0 * fy - 16.0);
  const float aval = (float)(500.0 * (fx - fy));
  const float bval = (float)(200.0 * (fy - fz));
  lab[3 * t] = lval / (-100);
  lab[3 * t + 1] = aval / 100;
  lab[3 * t + 2] = bval / 100;
}

__device__ __forceinline__ double lab_finv(double f) {
  const double c = (double)powf((float)f, 3.0f);
  return c > 0.008856 ? c : (f - 16.0 / 116.0) / 7.787;
}

__device__ __forceinline__ double linear_to_srgb(double v) {
  return v > 0.0031308 ? 1.055 * (double)powf((float)v, (float)(1.0 / 2.4)) - 0.055 : 12.92 * v;
}

__global__ void k_lab_to_rgb(const float* __restrict__ lab, float* __restrict__ rgb, long long n) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double L = lab[3 * t] * (-100), La = lab[3 * t + 1] * 100, Lb = lab[3 * t + 2] * 100;
  if (L != L || La != La || Lb != Lb) return;
  const double fy = (L + 16) / 116, fx = La / 500 + fy, fz = fy - Lb / 200;
  const double x = lab_finv(fx), y = lab_finv(fy), z = lab_finv(fz);
  const double X = 0.950456 * x, Y = 1.000 * y, Z = 1.088754 * z;
  const double R = X * 3.2406 + Y * (-1.5372) + Z * (-0.4986);
  const double G = X * -0.9689 + Y * 1.8758 + Z * 0.0415;
  const double B = X * 0.0557 + Y * (-0.2040) + Z * 1.0570;
  const double r = linear_to_srgb(R), g = linear_to_srgb(G), b = linear_to_srgb(B);
  float pr = (float)fmin(1.0, r * 1.0), pg = (float)fmin(1.0, g * 1.0), pb = (float)fmin(1.0, b * 1.0);
  pr = (float)fmax(0.0, (double)pr);
  pg = (float)fmax(0.0, (double)pg);
  pb = (float)fmax(0.0, (double)pb);
  rgb[3 * t] = pr;
  rgb[3 * t + 1] = pg;
  rgb[3 * t + 2] = pb;
}

void launch_rescale_u8(const uint8_t* in, float* out, long long n, float scale, hipStream_t s) {
  if (n <= 0) return;
  k_rescale_u8<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(in, out, n, scale);
}

void launch_rgb_to_lab(const float* bgr, float* lab, long long npix, hipStream_t s) {
  if (npix <= 0) return;
  k_rgb_to_lab<<<(unsigned)((npix + kBlock - 1) / kBlock), kBlock, 0, s>>>(bgr, lab, npix);
}

void launch_lab_to_rgb(const float* lab, float* rgb, long long npix, hipStream_t s) {
  if (npix <= 0) return;
  k_lab_to_rgb<<<(unsigned)((npix + kBlock - 1) / kBlock), kBlock, 0, s>>>(lab, rgb, npix);
}

}  // namespace stsp
