// Hand-written adjoint of the auto power spectrum (no counterpart in the
// reference, which has no gradients).
//
// Forward: P_b = sum_{k in H, bin(k) = b} h_k g_k |F_k|^2 / (M^2 C_b), with
// F = r2c(rho), H the half spectrum, h_k the Hermitian weight, g_k the CIC
// deconvolution factor and C_b the mode count.  For lambda_b = dL/dP_b:
//   A_k      = 2 lambda_{b(k)} g_k F_k / (M^2 C_{b(k)}),  A_0 = 0
//   dL/drho  = c2r(A)   (unnormalised inverse; interior kz implicitly doubled)
//   dL/dm_j  = sum_c w_jc G(c)
//   dL/dx_ja = m_j (d/box) sum_c (dw_jc/du_a) G(c)
// where w_jc are the CIC weights of fieldize.cpp:77-84 and G = dL/drho.
#include "fft_core.h"
#include "plan.h"

namespace genpk {

namespace {

__device__ __forceinline__ int64_t kv(int64_t n, int64_t d) { return n <= d / 2 ? n : n - d; }

template <typename T>
__global__ void k_weight(typename Cx<T>::C *__restrict__ f, int64_t d, int64_t hz, const int *__restrict__ q2bin,
                         const int *__restrict__ count, const double *__restrict__ lambda,
                         const float *__restrict__ wtab, double inv_m2) {
  using C = typename Cx<T>::C;
  const int64_t total = d * d * hz;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kz = i % hz;
    const int64_t row = i / hz;
    const int64_t x = row / d, y = row % d;
    const int64_t kx = kv(x, d), ky = kv(y, d);
    const int64_t q = kx * kx + ky * ky + kz * kz;
    C out = C{0, 0};
    if (q != 0) {
      const int b = q2bin[q];
      const int c = count[b];
      const float gw = (wtab[x] * wtab[y]) * wtab[kz];
      double g = (double)gw * (double)gw;
      g *= g;
      const double s = c ? 2.0 * lambda[b] * g * inv_m2 / (double)c : 0.0;
      const C v = f[i];
      out = C{(T)((double)v.x * s), (T)((double)v.y * s)};
    }
    f[i] = out;
  }
}

template <typename T, bool MASSES>
__global__ void k_gather(const T *__restrict__ G, int64_t d, int64_t rowlen, double units, int64_t n,
                         const float *__restrict__ pos, const float *__restrict__ masses, double mass,
                         float *__restrict__ gpos, float *__restrict__ gmass) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    int64_t i0[3], i1[3];
    double f[3];
    for (int a = 0; a < 3; a++) {
      const double u = (double)pos[3 * j + a] * units;
      const double fl = floor(u);
      f[a] = u - fl;
      long long c = (long long)fl % d;
      if (c < 0) c += d;
      i0[a] = c;
      i1[a] = (c + 1) % d;
    }
    const double m = MASSES ? (double)masses[j] : mass;
    double gm = 0, gx = 0, gy = 0, gz = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int bx = c & 1, by = (c >> 1) & 1, bz = (c >> 2) & 1;
      const double wx = bx ? f[0] : 1.0 - f[0];
      const double wy = by ? f[1] : 1.0 - f[1];
      const double wz = bz ? f[2] : 1.0 - f[2];
      const int64_t at = ((bx ? i1[0] : i0[0]) * d + (by ? i1[1] : i0[1])) * rowlen + (bz ? i1[2] : i0[2]);
      const double g = (double)G[at];
      gm += wx * wy * wz * g;
      gx += (bx ? 1.0 : -1.0) * wy * wz * g;
      gy += (by ? 1.0 : -1.0) * wx * wz * g;
      gz += (bz ? 1.0 : -1.0) * wx * wy * g;
    }
    const double s = m * units;
    gpos[3 * j + 0] = (float)(gx * s);
    gpos[3 * j + 1] = (float)(gy * s);
    gpos[3 * j + 2] = (float)(gz * s);
    if (gmass) gmass[j] = (float)gm;
  }
}

}  // namespace

void adjoint_weight(Plan &p, int nrbins, const double *lambda, double m) {
  GENPK_REQUIRE(p.fstate == FieldState::Spectrum, "adjoint needs the forward spectrum in the plan");
  GENPK_REQUIRE(m != 0, "total mass must be nonzero");
  const BinTables &bt = bin_tables(p, nrbins);
  const int64_t total = p.d * p.d * p.hz;
  const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(total, 256), 65536);
  {
  Stage st(p, "adjoint_weight");
  if (p.prec == 4)
    hipLaunchKernelGGL(k_weight<float>, dim3(grid), dim3(256), 0, p.stream, p.field.as<float2>(), p.d, p.hz,
                       bt.q2bin.as<int>(), bt.count.as<int>(), lambda, p.wtab.as<float>(), 1.0 / (m * m));
  else
    hipLaunchKernelGGL(k_weight<double>, dim3(grid), dim3(256), 0, p.stream, p.field.as<double2>(), p.d, p.hz,
                       bt.q2bin.as<int>(), bt.count.as<int>(), lambda, p.wtab.as<float>(), 1.0 / (m * m));
  GENPK_LAUNCH_CHECK();
  }
  c2r(p);
}

void adjoint_gather(Plan &p, double boxsize, int64_t n, const float *pos, const float *masses, double mass,
                    float *grad_pos, float *grad_mass) {
  GENPK_REQUIRE(p.fstate == FieldState::Real, "gather needs dL/drho in the plan (run the adjoint first)");
  GENPK_REQUIRE(boxsize > 0, "boxsize must be positive");
  if (n == 0) return;
  GENPK_REQUIRE(pos && grad_pos, "null positions or gradient output");
  const double units = (double)p.d / boxsize;
  const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(n, 256), 1 << 20);
  Stage st(p, "adjoint_gather");
  if (p.prec == 4) {
    if (masses)
      hipLaunchKernelGGL((k_gather<float, true>), dim3(grid), dim3(256), 0, p.stream, p.field.as<float>(), p.d,
                         p.rowlen, units, n, pos, masses, mass, grad_pos, grad_mass);
    else
      hipLaunchKernelGGL((k_gather<float, false>), dim3(grid), dim3(256), 0, p.stream, p.field.as<float>(), p.d,
                         p.rowlen, units, n, pos, masses, mass, grad_pos, grad_mass);
  } else {
    if (masses)
      hipLaunchKernelGGL((k_gather<double, true>), dim3(grid), dim3(256), 0, p.stream, p.field.as<double>(), p.d,
                         p.rowlen, units, n, pos, masses, mass, grad_pos, grad_mass);
    else
      hipLaunchKernelGGL((k_gather<double, false>), dim3(grid), dim3(256), 0, p.stream, p.field.as<double>(), p.d,
                         p.rowlen, units, n, pos, masses, mass, grad_pos, grad_mass);
  }
  GENPK_LAUNCH_CHECK();
}

}  // namespace genpk
