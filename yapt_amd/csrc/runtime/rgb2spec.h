// RGB -> sigmoid-polynomial spectrum coefficients (host).
//
// Restates SpectralUtility.cpp:253-609: for every node (r, g, b) = (x, y, z) / 63 of a 64^3 grid a
// Gauss-Newton solve (16 iterations, central-difference Jacobian, residual in the reference's CIELAB
// form) finds coefficients c such that sum_lambda sigmoid(c0 l^2 + c1 l + c2) * rgb(lambda) reproduces
// the node colour under D65. The renderer samples that grid with trilinear filtering.
//
// Nodes are solved on demand and cached, so converting the handful of material colours of a scene
// costs 8 solves per colour instead of the whole grid; `full_table` solves every node (needed only
// for the environment colour, which varies per ray).
#pragma once
#include <cmath>
#include <cstdint>
#include <mutex>
#include <unordered_map>
#include <vector>
#include <thread>
#include <algorithm>
#include <array>

namespace ypt {

class Rgb2Spec {
 public:
  static constexpr int RES = 64;
  static constexpr int CIE_RES = 321;

  // cie_xyz4: 321 x 4 floats (380..700 nm), d65: 321 floats, xyz_to_rgb / rgb_to_xyz: row-major 3x3
  void init(const double* cie_x, const double* cie_y, const double* cie_z, const double* d65,
            const double* xyz_to_rgb, const double* rgb_to_xyz) {
    double d65YSum = 0.0;
    for (int i = 0; i < CIE_RES; ++i) d65YSum += cie_y[i] * d65[i];
    double inv = 1.0 / d65YSum;
    conv_.resize(CIE_RES * 3);
    for (int i = 0; i < CIE_RES; ++i) {
      double X = cie_x[i] * d65[i] * inv, Y = cie_y[i] * d65[i] * inv, Z = cie_z[i] * d65[i] * inv;
      for (int r = 0; r < 3; ++r)
        conv_[i * 3 + r] = xyz_to_rgb[r * 3 + 0] * X + xyz_to_rgb[r * 3 + 1] * Y + xyz_to_rgb[r * 3 + 2] * Z;
    }
    for (int k = 0; k < 9; ++k) toXyz_[k] = rgb_to_xyz[k];
  }

  // coefficients of grid node (x, y, z) as float (the table stores float)
  void node(int x, int y, int z, float out[3]) {
    int key = (z * RES + y) * RES + x;
    {
      std::lock_guard<std::mutex> g(m_);
      auto it = cache_.find(key);
      if (it != cache_.end()) { out[0] = it->second[0]; out[1] = it->second[1]; out[2] = it->second[2]; return; }
    }
    double rgb[3] = {double(x) / (RES - 1), double(y) / (RES - 1), double(z) / (RES - 1)};
    double c[3] = {0, 0, 0};
    gauss_newton(rgb, c);
    std::array<float, 3> v{{(float)c[0], (float)c[1], (float)c[2]}};
    std::lock_guard<std::mutex> g(m_);
    cache_[key] = v;
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
  }

  // trilinear lookup with the half-texel remap, in FP32 (same arithmetic as the device lookup)
  void coeffs(const float rgb[3], float out[3]) {
    int i0[3], i1[3];
    float f[3];
    for (int a = 0; a < 3; ++a) {
      float u = rgb[a];
      float scale = (float)(RES - 1) / (float)RES, bias = 0.5f / (float)RES;
      float x = (u * scale + bias) * (float)RES - 0.5f;
      x = std::min(std::max(x, 0.f), (float)(RES - 1));
      float fl = std::floor(x);
      i0[a] = (int)fl;
      i1[a] = i0[a] + 1 < RES ? i0[a] + 1 : RES - 1;
      f[a] = x - fl;
      if (f[a] == 0.f) i1[a] = i0[a];  // a + 0 * (b - a) == a for the finite node values: skip the solve of b
    }
    float corner[2][2][2][3];
    for (int zz = 0; zz < 2; ++zz)
      for (int yy = 0; yy < 2; ++yy)
        for (int xx = 0; xx < 2; ++xx)
          node(xx ? i1[0] : i0[0], yy ? i1[1] : i0[1], zz ? i1[2] : i0[2], corner[zz][yy][xx]);
    for (int k = 0; k < 3; ++k) {
      float cz[2];
      for (int zz = 0; zz < 2; ++zz) {
        float cy[2];
        for (int yy = 0; yy < 2; ++yy) {
          float a = corner[zz][yy][0][k], b = corner[zz][yy][1][k];
          cy[yy] = a + f[0] * (b - a);
        }
        cz[zz] = cy[0] + f[1] * (cy[1] - cy[0]);
      }
      out[k] = cz[0] + f[2] * (cz[1] - cz[0]);
    }
  }

  // whole grid as RES^3 x 4 floats ([b][g][r], w = 0)
  void full_table(std::vector<float>& out, unsigned threads) {
    out.assign((size_t)RES * RES * RES * 4, 0.f);
    threads = std::max(1u, threads);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < threads; ++t)
      th.emplace_back([&, t]() {
        for (int z = (int)t; z < RES; z += (int)threads)
          for (int y = 0; y < RES; ++y)
            for (int x = 0; x < RES; ++x) {
              float c[3];
              node(x, y, z, c);
              size_t o = (((size_t)z * RES + y) * RES + x) * 4;
              out[o] = c[0]; out[o + 1] = c[1]; out[o + 2] = c[2];
            }
      });
    for (auto& x : th) x.join();
  }

  // sum_lambda sigmoid(poly) * rgb(lambda): the colour a coefficient triplet reproduces
  void reproduce(const double c[3], double rgb[3]) const { estimate(c, rgb); }

 private:
  std::vector<double> conv_;
  double toXyz_[9];
  std::mutex m_;
  std::unordered_map<int, std::array<float, 3>> cache_;

  static double sigm(double x) { return 0.5 * x / std::sqrt(1.0 + x * x) + 0.5; }

  void estimate(const double c[3], double out[3]) const {
    out[0] = out[1] = out[2] = 0.0;
    for (int i = 0; i < CIE_RES; ++i) {
      double lambda = double(i) / (CIE_RES - 1);
      double p = 0;
      for (int k = 0; k < 3; ++k) p = p * lambda + c[k];
      double s = sigm(p);
      out[0] += s * conv_[i * 3 + 0];
      out[1] += s * conv_[i * 3 + 1];
      out[2] += s * conv_[i * 3 + 2];
    }
  }
  void lab(const double rgb[3], double out[3]) const {
    const float Xn = 95.047f, Yn = 100.0f, Zn = 108.883f;
    const float delta = (float)(6.0 / 29.0);
    const float delta3 = delta * delta * delta;
    double xyz[3];
    for (int r = 0; r < 3; ++r) xyz[r] = toXyz_[r * 3 + 0] * rgb[0] + toXyz_[r * 3 + 1] * rgb[1] + toXyz_[r * 3 + 2] * rgb[2];
    double n[3] = {xyz[0] / (double)Xn, xyz[1] / (double)Yn, xyz[2] / (double)Zn};
    double ft[3];
    for (int i = 0; i < 3; ++i)
      ft[i] = n[i] > (double)delta3 ? std::pow(n[i], 1.0 / 3.0) : n[i] / (3.0 * delta * delta) + 4.0 / 29.0;
    out[0] = 116.0 * ft[1] - 16;
    out[1] = 500 * (ft[0] - ft[1]);
    out[2] = 200 * (ft[1] - ft[2]);
  }
  void residual(const double expected[3], const double c[3], double r[3]) const {
    double est[3], a[3], b[3];
    estimate(c, est);
    lab(expected, a);
    lab(est, b);
    for (int k = 0; k < 3; ++k) r[k] = a[k] - b[k];
  }
  // glm::inverse(dmat3) * v, with J[col][row]
  static void solve(const double J[3][3], const double v[3], double out[3]) {
    const double (*m)[3] = J;
    double det = +m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                 m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
    double o = 1.0 / det;
    double inv[3][3];
    inv[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]) * o;
    inv[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) * o;
    inv[2][0] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]) * o;
    inv[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) * o;
    inv[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]) * o;
    inv[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) * o;
    inv[0][2] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]) * o;
    inv[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) * o;
    inv[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]) * o;
    for (int r = 0; r < 3; ++r) out[r] = inv[0][r] * v[0] + inv[1][r] * v[1] + inv[2][r] * v[2];
  }
  void gauss_newton(const double expected[3], double c[3]) const {
    const double EPS = 1e-4;
    for (int it = 0; it < 16; ++it) {
      double r[3];
      residual(expected, c, r);
      double J[3][3];
      for (int i = 0; i < 3; ++i) {
        double cm[3] = {c[0], c[1], c[2]}, cp[3] = {c[0], c[1], c[2]};
        cm[i] -= EPS; cp[i] += EPS;
        double r0[3], r1[3];
        residual(expected, cm, r0);
        residual(expected, cp, r1);
        for (int k = 0; k < 3; ++k) J[i][k] = (r1[k] - r0[k]) * (1.0 / (2.0 * EPS));
      }
      double d[3];
      solve(J, r, d);
      for (int k = 0; k < 3; ++k) c[k] -= d[k];
      double err = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      if (err < 1e-6) break;
      double mx = std::max(std::max(c[0], c[1]), c[2]);
      if (mx > 200) for (int k = 0; k < 3; ++k) c[k] *= 200 / mx;
    }
  }
};

}  // namespace ypt
