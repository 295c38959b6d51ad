// stage_args.h -- POD launch descriptor shared by the host code and the
// stage kernels (passed by value as the kernel argument).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gpe {

// One fused sweep of the time-stepping path.  Each kind reads psi and the
// k-fields its stage input needs, evaluates f(u) on the 2nd/4th-order
// stencil and writes what the next stage (or the new state) needs.
enum Kind : int {
  K_RHS = 0,    // out = f(psi)                       (get_rhs, solve.f90:382)
  K_EULER,      // psi' = psi + dt f(psi)              (solve.f90:57-74)
  K_RK2_1,      // k1 = f(psi)                         (solve.f90:91)
  K_RK2_2,      // psi' = psi + dt f(psi + dt/2 k1)     (solve.f90:92-99)
  K_RK4_1,      // k1 = f(psi)                         (solve.f90:124)
  K_RK4_2,      // k2 = f(psi + dt/2 k1); acc = k1/6 + k2/3
  K_RK4_3,      // k3 = f(psi + dt/2 k2)
  K_RK4_4,      // psi' = psi + dt((acc + k3/3) + f(psi + dt k3)/6)  (144-148)
  K_CK_1,       // kk1 = dt f(psi); scale = |psi| + dt|f| + 1e-30 (176-178, 293-296)
  K_CK_2, K_CK_3, K_CK_4, K_CK_5, // kk_s = dt f(psi + sum b_sm kk_m)
  K_CK_6,       // kk6; psi' = psi + sum c kk; err = sum dc kk; max|err/scale|
  K_ENERGY,     // sum Re(-lap conj psi) + 0.5|psi|^4  (variables.f90:580-604)
  K_COUNT
};

// Cash-Karp tableau (solve.f90:255-281), each weight formed in double as
// the reference forms its real(pr) parameters.  The kernels take them as
// compile-time constants: no scalar registers hold them across the z march.
// ck_b(s, m): coefficient of kk_{m+1} in the input of stage s (2..6).
__host__ __device__ constexpr double ck_b(int s, int m) {
  return s == 2   ? (m == 0 ? 0.2 : 0.0)
         : s == 3 ? (m == 0 ? 0.075 : m == 1 ? 0.225 : 0.0)
         : s == 4 ? (m == 0 ? 0.3 : m == 1 ? -0.9 : m == 2 ? 1.2 : 0.0)
         : s == 5 ? (m == 0 ? -11.0 / 54.0 : m == 1 ? 2.5 : m == 2 ? -70.0 / 27.0 : m == 3 ? 35.0 / 27.0 : 0.0)
         : s == 6 ? (m == 0   ? 1631.0 / 55296.0
                     : m == 1 ? 175.0 / 512.0
                     : m == 2 ? 575.0 / 13824.0
                     : m == 3 ? 44275.0 / 110592.0
                     : m == 4 ? 253.0 / 4096.0
                              : 0.0)
                  : 0.0;
}
// ck_c(m): weight of kk_m in the fifth-order solution (m = 1..6).
__host__ __device__ constexpr double ck_c(int m) {
  return m == 1 ? 37.0 / 378.0 : m == 3 ? 250.0 / 621.0 : m == 4 ? 125.0 / 594.0 : m == 6 ? 512.0 / 1771.0 : 0.0;
}
// ck_dc(m): c_m minus the embedded fourth-order weight (the error estimate).
__host__ __device__ constexpr double ck_dc(int m) {
  return m == 1   ? ck_c(1) - 2825.0 / 27648.0
         : m == 2 ? ck_c(2) - 0.0
         : m == 3 ? ck_c(3) - 18575.0 / 48384.0
         : m == 4 ? ck_c(4) - 13525.0 / 55296.0
         : m == 5 ? ck_c(5) - 277.0 / 14336.0
         : m == 6 ? ck_c(6) - 0.25
                  : 0.0;
}

// Two consecutive stages fused into one sweep (fused.hpp).
enum Pair : int { P_RK4_12 = 0, P_RK4_34, P_CK_12, P_CK_34, P_CK_56, P_COUNT };

// Division by a real constant as the reference build evaluates it: the
// divisor is promoted to complex and divided by compiler-rt's logb-scaled
// quotient, i.e. ((a*cs + b*0)/den, (b*cs - a*0)/den) * 2^-e with
// cs = s*2^-e, den = cs*cs + 0*0, e = logb(s).
struct DivC {
  double cs, den;
  int e;
  int pad;
};

struct StageArgs {
  // geometry and tiling
  int nx, ny, nz;
  int periodic;  // bcs == 1
  int eqn;
  int ntx, nty, nzc, zc;
  int nblocks;
  int write_scale;  // K_CK_1 on the first attempt of a step
  int use_diss;     // diss_amp != 0
  int use_urhs;     // urhs != 0
  int scale_real;   // rk45 error scale stored as one real word (real time)
  // coordinates and coefficients (ic.f90:32-69, 713-731)
  double xl, yl, zl, dx, dy, dz;
  double g, mu, ox2, oy2, oz2, urhs, diss_amp, a_diss, b_diss, c_diss;
  // FP64 divisions in the reference's form
  DivC d_xx, d_yy, d_zz;  // dx2 (order 2) or 12*dx2 (order 4)
  DivC d_x;               // 2*dx or 12*dx
  DivC d6, d3;            // RK4 combination /6 and /3
  // FP32 reciprocals
  double inv_xx, inv_yy, inv_zz, inv_x;
  // time step (complex dt) and stage coefficients
  double dtr, dti;   // dt
  double hr, hi;     // 0.5*dt
  double sdtr, sdti; // dt at entry of the rk45 step (scale)
  double inv_absdt;  // 1/|dt| (compensated rk45: |f| = |kk1|/|dt|)
  // norm weights (variables.f90:49-51)
  double w1, w2, w3;
  // fields (device pointers; complex interleaved)
  const void *psi;
  const void *lo;          // FP32C: low part of psi
  const void *kin[5];      // input k-fields
  const void *acc_in;      // K_RK4_4
  const void *scale_in;    // K_CK_6
  void *kout;              // k / kk output
  void *acc_out;           // K_RK4_2
  void *psi_out;           // new state
  void *lo_out;            // FP32C new low part
  void *scale_out;         // K_CK_1
  double *part_norm;       // per-block partials
  double *part_emax;
  double *part_energy;
};

// A tile of one instance in a batched sweep.
struct WorkItem {
  int inst;    // instance slot (index into the kind's StageArgs table)
  int tx, ty;  // 64 x TY column tile
  int z0, z1;  // planes z0 .. z1-1
  int slot;    // partial-sum slot: contiguous per instance, in tile order
};

// Per-instance step state of a batch (device memory; see batch_ctl.hpp).
struct InstCtl {
  int active;       // 1 while steps remain
  int status;       // 0 running, 1 done, -1 NaN in the error norm, -2 too many rejections
  int scheme;       // 1, 2, 4 or 45
  int real_time;
  int cur;          // which of buf[0], buf[1] holds the state
  int first_attempt;
  int n_rejects;    // rejections in the current rk45 step
  int last_rejects; // rejections of the last accepted step
  long long steps_done, steps_target;
  int k0, nk;       // stage kinds of the scheme: args[(k0 .. k0+nk-1) * cap + slot]
  double t, im_t, dt_re, dt_im, sdt_re, sdt_im;
  double errmax, norm;
  double safety, eps, dt_inc, dt_dec, errcon, dv;
  void *buf[2], *lob[2];
};

// A zero point of the vortex-line length pass (see diag.hpp).
struct LinePoint {
  long long cell;  // linear cell index, x fastest
  double x, y, z;
};

// A vortex line to imprint (type line_param, parameters.in:33-43) on a
// grid, with the reference's safe wavelengths already applied.
struct LineArgs {
  double xl, yl, zl, dx, dy, dz;
  double x0, y0, z0, a1, a2, ll, sgn, scal, safe1, safe2;
  int nx, ny, nz, dir, phase_only;
};

// A vortex-core point (zeros pass, see diag.hpp); the layout of
// gpe_zero_point in include/gpe_hip.h.
struct ZeroPoint {
  int section;     // section of the reference's output file, from 0
  int reserved;
  long long cell;  // i + nx*(j + ny*k) of the lower corner
  double x, y, z;
};

}  // namespace gpe
