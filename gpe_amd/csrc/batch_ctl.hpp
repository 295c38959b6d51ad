// batch_ctl.hpp -- per-instance step control of a batch, on the device.
//
// After the final stage of every batched iteration, one wave per instance
// reduces that instance's partial sums (in its own slot order, so the result
// does not depend on what else is in the batch) and advances its state as
// the reference's integrators do: rk4 (solve.f90:150-151), or the
// accept/reject and step-size control of rkqs (solve.f90:180-229).  It then
// writes the new dt and the swapped state buffers into the instance's
// StageArgs for every stage kind, so the captured step graph can be replayed
// without a host round trip.
#pragma once
#include "stencil.hpp"

namespace gpe {


// StageArgs laid out args[kind * cap + inst]; range[2i], range[2i+1] are
// the partial slots [begin, end) of instance i in this iteration's final
// stage.
__global__ void __launch_bounds__(64) batch_control(InstCtl *__restrict__ ctl, StageArgs *__restrict__ args,
                                                     int cap, int *__restrict__ active, const int *__restrict__ range,
                                                     const double *__restrict__ pn, const double *__restrict__ pe) {
  const int i = blockIdx.x;
  InstCtl &C = ctl[i];
  if (!C.active) return;
  const int lane = threadIdx.x;
  double sn = 0.0, mx = 0.0;
  const int sb = range[2 * i], se = range[2 * i + 1];
  for (int s = sb + lane; s < se; s += 64) {
    sn += pn[s];
    if (pe) mx = (pe[s] > mx || pe[s] != pe[s]) ? pe[s] : mx;
  }
  for (int o = 32; o > 0; o >>= 1) {
    sn += __shfl_xor(sn, o);
    const double om = __shfl_xor(mx, o);
    mx = (om > mx || om != om) ? om : mx;
  }
  if (lane != 0) return;

  bool accepted = true;
  if (C.scheme == 45) {
    const double em = mx / C.eps;  // maxval(abs(err/psi_scal))/eps
    if (em != em) {
      C.status = -1;
      C.active = 0;
      active[i] = 0;
      return;
    }
    if (em <= 1.0) {
      double nr, ni;
      if (em > C.errcon) {
        const double pw = refm::pow_cr(em, C.dt_inc);
        nr = (C.safety * C.dt_re) * pw;
        ni = (C.safety * C.dt_im) * pw;
      } else {
        nr = 5.0 * C.dt_re;
        ni = 5.0 * C.dt_im;
      }
      C.t = C.t + C.dt_re;
      C.dt_re = nr;
      C.dt_im = ni;
      C.im_t = C.im_t + fabs(ni);
      C.errmax = em;
      C.last_rejects = C.n_rejects;
      C.n_rejects = 0;
    } else {
      accepted = false;
      if (++C.n_rejects > 1000) {
        C.status = -2;
        C.active = 0;
        active[i] = 0;
        return;
      }
      const double pw = refm::pow_cr(em, C.dt_dec);
      const double ar = (C.safety * C.dt_re) * pw, ai = (C.safety * C.dt_im) * pw;
      const double a = refm::hypot_ref(ar, ai), b = 0.1 * refm::hypot_ref(C.dt_re, C.dt_im);
      const double m = a > b ? a : b;
      if (C.real_time) {
        C.dt_re = m;
        C.dt_im = 0.0;
      } else {
        C.dt_re = 0.0;
        C.dt_im = -m;
      }
    }
  } else {
    C.t = C.t + C.dt_re;
    C.im_t = C.im_t + fabs(C.dt_im);
  }
  if (accepted) {
    C.norm = sn * C.dv;
    C.cur ^= 1;
    C.steps_done += 1;
    C.first_attempt = 1;
    C.sdt_re = C.dt_re;
    C.sdt_im = C.dt_im;
    if (C.steps_done >= C.steps_target) {
      C.active = 0;
      C.status = 1;
      active[i] = 0;
    }
  } else {
    C.first_attempt = 0;
  }
  for (int k = C.k0; k < C.k0 + C.nk; ++k) {
    StageArgs &A = args[k * cap + i];
    A.dtr = C.dt_re;
    A.dti = C.dt_im;
    A.hr = 0.5 * C.dt_re;
    A.hi = 0.5 * C.dt_im;
    A.sdtr = C.sdt_re;
    A.sdti = C.sdt_im;
    A.inv_absdt = 1.0 / refm::hypot_ref(C.dt_re, C.dt_im);
    A.write_scale = C.first_attempt;
    A.psi = C.buf[C.cur];
    A.psi_out = C.buf[C.cur ^ 1];
    A.lo = C.lob[C.cur];
    A.lo_out = C.lob[C.cur ^ 1];
  }
}

}  // namespace gpe
