// fused.hpp -- two consecutive stages of an integrator in one sweep.
//
// Stage B's input u_B = w + b*k_A needs stage A's output k_A at B's stencil
// points, so a workgroup computes k_A on its whole 64 x TY tile and k_B on
// the tile's interior (64-2R) x (TY-2R); tiles overlap by 2R columns and
// rows.  Along z the sweep runs stage A R planes ahead of stage B: at step
// k it loads plane k+2R, computes k_A at plane k+R (z-neighbours from a
// register queue of u_A, x/y from LDS tile A with its ghost ring) and k_B at
// plane k (z-neighbours from a register queue of u_B, x/y from LDS tile B).
// Pairs: RK4 stages (1,2) and (3,4); Cash-Karp stages (1,2), (3,4), (5,6).
// Fields moved per site and sweep (passes), against the stages one by one:
//   RK4 (1,2): read psi; write k2, acc                  3  (2 + 4)
//   RK4 (3,4): read psi, k2, acc; write psi'            4  (3 + 4)
//   CK (1,2):  read psi; write kk1, kk2                 3  (2 + 3)
//   CK (3,4):  read psi, kk1, kk2; write kk3, kk4       5  (4 + 5)
//   CK (5,6):  read psi, kk1..kk4; write psi'           6  (6 + 7)
// (RK4: 7 passes per step instead of 13; Cash-Karp: 14 per attempt instead
// of 27; k1, k3 and kk5 are consumed inside the sweep and never stored.)
//
// Every value is computed by the same expressions, in the same order, as
// the one-stage kernels (stage_tile), so the fields are bitwise identical
// to the unfused path.  Boundaries: tile points outside the grid take the
// index-mapped site, as the one-stage kernels' loads do.  With periodic
// walls a point outside the grid is the wrapped site with its neighbours in
// the same orientation, so its k_A is computed in place.  With reflective
// walls the mirror image reverses the neighbour order, which the
// reference's non-commutative stencil sums would round differently, so
// such points take u_B from their mirror partner inside the tile: x/y
// through redirected LDS reads, z by copying queue entries.
#pragma once
#include "stencil.hpp"

namespace gpe {

template <int PAIR>
struct PairInfo {
  static constexpr int KA = PAIR == P_RK4_12   ? K_RK4_1
                            : PAIR == P_RK4_34 ? K_RK4_3
                            : PAIR == P_CK_12  ? K_CK_1
                            : PAIR == P_CK_34  ? K_CK_3
                                               : K_CK_5;
  static constexpr int KB = KA + 1;
  static constexpr bool store_a = PAIR == P_CK_12 || PAIR == P_CK_34;
  static constexpr bool is_ck = PAIR >= P_CK_12;
};

template <typename T, int R, int NW, int P, int D, int PAIR, bool COMP>
__device__ __forceinline__ void pair_tile(const StageArgs &A, const StageArgs &B, const int tx, const int ty,
                                          const int z0, const int z1, const int slot, double *pn, double *pe) {
  using PI = PairInfo<PAIR>;
  constexpr int KA = PI::KA, KB = PI::KB;
  constexpr int TX = 64, TY = NW * P, TXO = TX - 2 * R, TYO = TY - 2 * R;
  constexpr int LXA = TX + 2 * R, LYA = TY + 2 * R;
  constexpr int NQ = 2 * R + 1;
  constexpr int NT = 64 * NW;
  constexpr int NIN = KindInfo<KA>::nin;  // k-fields read with psi
  constexpr int NINA = NIN > 0 ? NIN : 1;
  constexpr bool HAS_ACC = KindInfo<KB>::has_acc;
  constexpr int NF = FifoInfo<KB, COMP>::nf;
  constexpr int NFA = NF > 0 ? NF : 1;
  constexpr int NO = OutInfo<KB, COMP>::no;
  constexpr int NOA = NO > 0 ? NO : 1;
  constexpr bool IS_DBL = std::is_same<T, double>::value;
  constexpr int NYH = 2 * R * TX, NXH = 2 * R * TY, NH = NYH + NXH;
  constexpr int HPT = (NH + NT - 1) / NT;
  using V = vec2_t<T>;
  static_assert(TY > 2 * R, "tile must have interior rows");

  __shared__ V ldsA[LYA * LXA];
  __shared__ V ldsB[TY * TX];
  __shared__ double red[2][NW];

  struct Raw {
    cx<T> ps;
    cx<T> kk[NINA];
    cx<T> ac;
  };

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nx = A.nx, ny = A.ny, nz = A.nz, per = A.periodic;
  const size_t plane = (size_t)nx * (size_t)ny;
  const int gx0 = tx * TXO - R, gy0 = ty * TYO - R;

  // own points: column gi (mapped gim), rows gj[p] (mapped gjm[p])
  const int gi = gx0 + lane;
  // (stage A is needed up to R points past the last site, and its stencil
  // reaches R further)
  const int gim = bcmap(min(gi, nx - 1 + 2 * R), nx, per);
  const bool out_x = lane >= R && lane < TX - R && gi < nx;
  int gj[P], gjm[P], roff[P];
  bool out_y[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int r = wave * P + p;
    gj[p] = gy0 + r;
    gjm[p] = bcmap(min(gj[p], ny - 1 + 2 * R), ny, per);
    roff[p] = gjm[p] * nx + gim;
    out_y[p] = r >= R && r < TY - R && gj[p] < ny;
  }
  // ghost ring of LDS tile A
  int hl[HPT], hoff[HPT];
#pragma unroll
  for (int i = 0; i < HPT; ++i) {
    const int h = threadIdx.x + i * NT;
    hl[i] = -1;
    hoff[i] = 0;
    if (h < NH) {
      int lr, lc;
      if (h < NYH) {
        const int rs = h / TX, c = h - rs * TX;
        lr = rs < R ? rs : TY + rs;
        lc = c + R;
      } else {
        const int hh = h - NYH;
        const int r = hh / (2 * R), cs = hh - r * (2 * R);
        lr = r + R;
        lc = cs < R ? cs : TX + cs;
      }
      hl[i] = lr * LXA + lc;
      const int hgi = bcmap(min(gx0 + lc - R, nx - 1 + 2 * R), nx, per);
      const int hgj = bcmap(min(gy0 + lr - R, ny - 1 + 2 * R), ny, per);
      hoff[i] = hgj * nx + hgi;
    }
  }
  // stage-B stencil reads in LDS tile B: offsets -R..R (without 0), with
  // reflective walls redirected to the mirror partner
  int bx[2 * R], by[P][2 * R];
#pragma unroll
  for (int d = 0; d < 2 * R; ++d) {
    const int o = d < R ? d - R : d - R + 1;
    int c = lane + o;
    const int g = gi + o;
    if (!per && (g < 0 || g > nx - 1)) c += bcmap(g, nx, 0) - g;
    bx[d] = min(max(c, 0), TX - 1);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int r0 = wave * P + p;
      int r = r0 + o;
      const int h = gj[p] + o;
      if (!per && (h < 0 || h > ny - 1)) r += bcmap(h, ny, 0) - h;
      by[p][d] = min(max(r, 0), TY - 1);
    }
  }

  // coefficients: A's stage input, B's stage input (w + bB * k_A)
  const cx<T> HA = mk((T)A.hr, (T)A.hi);  // RK4 stages 2 and 3
  const cx<T> HB = KB == K_RK4_4 ? mk((T)B.dtr, (T)B.dti) : mk((T)B.hr, (T)B.hi);

  auto load_raw = [&](size_t s, Raw &r, const bool full) {
    r.ps = ldc<T>(A.psi, s);
    if constexpr (NIN > 0) {
#pragma unroll
      for (int m = 0; m < NIN; ++m) r.kk[m] = ldc<T>(A.kin[m], s);
    }
    if constexpr (HAS_ACC) {
      if (full) r.ac = ldc<T>(B.acc_in, s);
    }
  };
  auto make_uA = [&](const Raw &r) -> cx<T> {
    if constexpr (NIN == 0) {
      return r.ps;
    } else if constexpr (KindInfo<KA>::complex_coef) {
      return r.ps + cmul(HA, r.kk[0]);
    } else {
      cx<T> u = r.ps;
#pragma unroll
      for (int m = 0; m < NIN; ++m) u = axpy<COMP>((T)ck_b(KA - K_CK_1 + 1, m), r.kk[m], u);
      return u;
    }
  };
  // u_B without its k_A term (same order of additions as the one-stage kernel)
  auto make_w = [&](const Raw &r) -> cx<T> {
    if constexpr (KindInfo<KB>::complex_coef) {
      return r.ps;
    } else {
      cx<T> u = r.ps;
#pragma unroll
      for (int m = 0; m < NIN; ++m) u = axpy<COMP>((T)ck_b(KB - K_CK_1 + 1, m), r.kk[m], u);
      return u;
    }
  };
  auto add_kA = [&](cx<T> w, cx<T> ka) -> cx<T> {
    if constexpr (KindInfo<KB>::complex_coef)
      return w + cmul(HB, ka);
    else
      return axpy<COMP>((T)ck_b(KB - K_CK_1 + 1, NIN), ka, w);
  };
  // B's final-combination partials known at load time
  auto make_F = [&](const Raw &r, cx<T> *F) {
    if constexpr (KB == K_RK4_4) {
      F[0] = r.ps;
      F[1] = r.ac;
    } else if constexpr (KB == K_CK_6) {
      cx<T> e = rmul((T)ck_dc(1), r.kk[0]);
#pragma unroll
      for (int m = 1; m < 4; ++m) e = axpy<COMP>((T)ck_dc(m + 1), r.kk[m], e);
      if constexpr (COMP) {
        cx<T> pd = rmul((T)ck_c(1), r.kk[0]);  // c2 = c5 = 0
        pd = axpy<COMP>((T)ck_c(3), r.kk[2], pd);
        pd = axpy<COMP>((T)ck_c(4), r.kk[3], pd);
        F[0] = r.ps;
        F[1] = pd;
        F[2] = e;
        F[3] = scale_from_kk1(B, r.ps, r.kk[0]);
      } else {
        cx<T> s = r.ps;
#pragma unroll
        for (int m = 0; m < 4; ++m) s = s + rmul((T)ck_c(m + 1), r.kk[m]);
        F[0] = s;
        F[1] = e;
      }
    }
  };
  // ... and the k_A terms added when k_A is known (last in each sum)
  auto finish_F = [&](cx<T> *F, cx<T> ka) {
    if constexpr (KB == K_RK4_2) {
      F[0] = rdiv(ka, B.d6, 1.0 / 6.0);
    } else if constexpr (KB == K_RK4_4) {
      F[1] = F[1] + rdiv(ka, B.d3, 1.0 / 3.0);
    } else if constexpr (KB == K_CK_6) {
      if constexpr (COMP) {
        F[2] = axpy<COMP>((T)ck_dc(5), ka, F[2]);
      } else {
        F[0] = F[0] + rmul((T)ck_c(5), ka);
        F[1] = F[1] + rmul((T)ck_dc(5), ka);
      }
    }
  };
  auto load_out = [&](size_t s, cx<T> *o) {
    int i = 0;
    if constexpr (KB == K_CK_6 && !COMP) {
      if (B.scale_real)
        o[i++] = mk(static_cast<const T *>(B.scale_in)[s], (T)0);
      else
        o[i++] = ldc<T>(B.scale_in, s);
    }
    if constexpr (COMP && KindInfo<KB>::final_) o[i++] = ldc<T>(B.lo, s);
    (void)i;
  };

  cx<T> qA[NQ][P];          // u_A, planes k .. k+2R
  cx<T> wF[R + 1][P];       // w, planes k+R .. k+2R
  cx<T> FF[NQ][P][NFA];     // B's partials, planes k .. k+2R
  cx<T> G[R + 1][HPT];      // u_A ghost ring, planes k+R .. k+2R
  cx<T> qB[NQ][P];          // u_B, planes k-R .. k+R
#pragma unroll
  for (int m = 0; m < NQ; ++m)
#pragma unroll
    for (int p = 0; p < P; ++p) {
      qB[m][p] = mk((T)0, (T)0);
#pragma unroll
      for (int f = 0; f < NFA; ++f) FF[m][p][f] = mk((T)0, (T)0);
    }

  const int kb = z0 - 2 * R;  // first step (stage A only until z0)
  // prime: u_A of planes kb .. kb+2R-1, w and ghosts of kb+R .. kb+2R-1
#pragma unroll
  for (int m = 0; m < 2 * R; ++m) {
    const size_t zoff = (size_t)bcmap(kb + m, nz, per) * plane;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      Raw r;
      load_raw(zoff + roff[p], r, false);
      qA[m][p] = make_uA(r);
      if (m >= R) wF[m - R][p] = make_w(r);
    }
    if (m >= R) {
#pragma unroll
      for (int i = 0; i < HPT; ++i) {
        if (hl[i] >= 0) {
          Raw r;
          load_raw(zoff + hoff[i], r, false);
          G[m - R][i] = make_uA(r);
        } else {
          G[m - R][i] = mk((T)0, (T)0);
        }
      }
    }
  }

  // D planes of loads in flight: a fused sweep reads few fields per plane
  // and its register footprint allows few workgroups per CU, so one plane
  // of prefetch would leave HBM idle between planes
  Raw rn[D][P];
  Raw hn[D][HPT];
  cx<T> on[P][NOA];
  auto issue_raw = [&](int k, Raw *rb, Raw *hb) {
    const size_t zi = (size_t)bcmap(k + 2 * R, nz, per) * plane;
#pragma unroll
    for (int p = 0; p < P; ++p) load_raw(zi + roff[p], rb[p], true);
#pragma unroll
    for (int i = 0; i < HPT; ++i)
      if (hl[i] >= 0) load_raw(zi + hoff[i], hb[i], false);
  };
  auto issue_out = [&](int k) {
    if constexpr (NO > 0) {
#pragma unroll
      for (int p = 0; p < P; ++p) load_out((size_t)k * plane + roff[p], on[p]);
    }
  };
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (kb + d < z1) issue_raw(kb + d, rn[d], hn[d]);

  double acc_norm = 0.0, acc_emax = 0.0;

  // one step of the march; rb/hb hold plane k+2R and are refilled with
  // plane k+D+2R
  auto step = [&](const int k, Raw *rb, Raw *hb) {
    // plane k+2R arrives
#pragma unroll
    for (int p = 0; p < P; ++p) {
      qA[2 * R][p] = make_uA(rb[p]);
      wF[R][p] = make_w(rb[p]);
      if constexpr (NF > 0) make_F(rb[p], FF[2 * R][p]);
    }
#pragma unroll
    for (int i = 0; i < HPT; ++i) G[R][i] = hl[i] >= 0 ? make_uA(hb[i]) : mk((T)0, (T)0);
    cx<T> oc[P][NOA];
    if constexpr (NO > 0) {
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int o = 0; o < NO; ++o) oc[p][o] = on[p][o];
    }
    if (k + D < z1) issue_raw(k + D, rb, hb);
    if (k + 1 < z1 && k + 1 >= z0) issue_out(k + 1);

    // ---- stage A at plane k+R ----
    const int ka = k + R;
    const int kam = bcmap(min(ka, nz - 1 + R), nz, per);
#pragma unroll
    for (int i = 0; i < HPT; ++i) {
      if (hl[i] >= 0) {
        V v;
        v.x = G[0][i].re;
        v.y = G[0][i].im;
        ldsA[hl[i]] = v;
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      V v;
      v.x = qA[R][p].re;
      v.y = qA[R][p].im;
      ldsA[(wave * P + p + R) * LXA + lane + R] = v;
    }
    lds_barrier();
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int c = (wave * P + p + R) * LXA + lane + R;
      auto LD = [&](int off) {
        V v = ldsA[c + off];
        return mk<T>(v.x, v.y);
      };
      const cx<T> c0 = qA[R][p];
      cx<T> xm1 = LD(-1), xp1 = LD(1), ym1 = LD(-LXA), yp1 = LD(LXA);
      cx<T> xm2 = xm1, xp2 = xp1, ym2 = ym1, yp2 = yp1;
      if constexpr (R == 2) {
        xm2 = LD(-2);
        xp2 = LD(2);
        ym2 = LD(-2 * LXA);
        yp2 = LD(2 * LXA);
      }
      const cx<T> dxx = d2<T, R>(xm2, xm1, c0, xp1, xp2, A.d_xx, A.inv_xx);
      const cx<T> dyy = d2<T, R>(ym2, ym1, c0, yp1, yp2, A.d_yy, A.inv_yy);
      const cx<T> dzz = d2<T, R>(qA[0][p], qA[R - 1][p], c0, qA[R + 1][p], qA[2 * R][p], A.d_zz, A.inv_zz);
      const cx<T> lap = (dxx + dyy) + dzz;
      cx<T> dpx = mk((T)0, (T)0);
      if (A.use_urhs) dpx = d1<T, R>(xm2, xm1, xp1, xp2, A.d_x, A.inv_x);
      const cx<T> f = rhs_site<T>(A, c0, lap, dpx, gim, gjm[p], kam);
      cx<T> kA;
      if constexpr (PI::is_ck) {
        if constexpr (IS_DBL || !COMP) {
          kA = cmul(mk((T)A.dtr, (T)A.dti), f);
        } else {
          cx<double> kd = cmul(mk(A.dtr, A.dti), todbl(f));
          kA = mk((T)kd.re, (T)kd.im);
        }
      } else {
        kA = f;
      }
      if constexpr (PI::store_a || KA == K_CK_1) {
        if (out_x && out_y[p] && ka >= z0 && ka < z1) {
          const size_t s = (size_t)ka * plane + (size_t)gj[p] * nx + gi;
          if constexpr (PI::store_a) stc<T>(A.kout, s, kA);
          if constexpr (KA == K_CK_1 && !COMP) {
            if (A.write_scale) {
              // psi_scal = abs(in) + dt*abs(f) + 1e-30 (solve.f90:176-178)
              T af = cabs_(f), au = cabs_(c0);
              cx<T> sc = mk((au + (T)A.sdtr * af) + (T)1e-30, (T)A.sdti * af);
              if (A.scale_real)
                static_cast<T *>(A.scale_out)[s] = sc.re;
              else
                stc<T>(A.scale_out, s, sc);
            }
          }
        }
      }
      qB[2 * R][p] = add_kA(wF[0][p], kA);
      if constexpr (NF > 0) finish_F(FF[R][p], kA);
    }
    if (!per) {
      // reflective walls: planes beyond the top take their mirror's u_B
#pragma unroll
      for (int j = 1; j <= R; ++j)
        if (ka == nz - 1 + j)
#pragma unroll
          for (int p = 0; p < P; ++p) qB[2 * R][p] = qB[2 * R - 2 * j][p];
      // ... and below the bottom, once their mirrors are in the queue
      if (k == 0) {
#pragma unroll
        for (int j = 1; j <= R; ++j)
#pragma unroll
          for (int p = 0; p < P; ++p) qB[R - j][p] = qB[R + j][p];
      }
    }

    // ---- stage B at plane k ----
    if (k >= z0) {
#pragma unroll
      for (int p = 0; p < P; ++p) {
        V v;
        v.x = qB[R][p].re;
        v.y = qB[R][p].im;
        ldsB[(wave * P + p) * TX + lane] = v;
      }
    }
    lds_barrier();  // also orders this step's reads of tile A before the next step's writes
    if (k >= z0) {
#pragma unroll
      for (int p = 0; p < P; ++p) {
        if (!(out_x && out_y[p])) continue;
        const int row = (wave * P + p) * TX;
        auto LX = [&](int d) {
          V v = ldsB[row + bx[d]];
          return mk<T>(v.x, v.y);
        };
        auto LY = [&](int d) {
          V v = ldsB[by[p][d] * TX + lane];
          return mk<T>(v.x, v.y);
        };
        const cx<T> c0 = qB[R][p];
        // bx/by order: offsets -R..-1, 1..R
        cx<T> xm1 = LX(R - 1), xp1 = LX(R), ym1 = LY(R - 1), yp1 = LY(R);
        cx<T> xm2 = xm1, xp2 = xp1, ym2 = ym1, yp2 = yp1;
        if constexpr (R == 2) {
          xm2 = LX(0);
          xp2 = LX(3);
          ym2 = LY(0);
          yp2 = LY(3);
        }
        const cx<T> dxx = d2<T, R>(xm2, xm1, c0, xp1, xp2, B.d_xx, B.inv_xx);
        const cx<T> dyy = d2<T, R>(ym2, ym1, c0, yp1, yp2, B.d_yy, B.inv_yy);
        const cx<T> dzz = d2<T, R>(qB[0][p], qB[R - 1][p], c0, qB[R + 1][p], qB[2 * R][p], B.d_zz, B.inv_zz);
        const cx<T> lap = (dxx + dyy) + dzz;
        cx<T> dpx = mk((T)0, (T)0);
        if (B.use_urhs) dpx = d1<T, R>(xm2, xm1, xp1, xp2, B.d_x, B.inv_x);
        const cx<T> f = rhs_site<T>(B, c0, lap, dpx, gi, gj[p], k);
        const size_t s = (size_t)k * plane + (size_t)gj[p] * nx + gi;
        cx<T> F[NFA];
#pragma unroll
        for (int q = 0; q < NFA; ++q) F[q] = FF[0][p][q];
        if constexpr (KB == K_RK4_2) {
          stc<T>(B.kout, s, f);
          stc<T>(B.acc_out, s, F[0] + rdiv(f, B.d3, 1.0 / 3.0));
        } else if constexpr (KB == K_RK4_4) {
          const cx<T> ps = F[0];
          const cx<T> incr = F[1] + rdiv(f, B.d6, 1.0 / 6.0);
          cx<double> np;
          if constexpr (IS_DBL) {
            np = ps + cmul(mk((T)B.dtr, (T)B.dti), incr);
            stc<T>(B.psi_out, s, np);
          } else if constexpr (COMP) {
            np = (todbl(ps) + todbl(oc[p][0])) + cmul(mk(B.dtr, B.dti), todbl(incr));
            const cx<T> hi = mk((T)np.re, (T)np.im);
            stc<T>(B.psi_out, s, hi);
            stc<T>(B.lo_out, s, mk((T)(np.re - (double)hi.re), (T)(np.im - (double)hi.im)));
          } else {
            const cx<T> nf = ps + cmul(mk((T)B.dtr, (T)B.dti), incr);
            stc<T>(B.psi_out, s, nf);
            np = todbl(nf);
          }
          const double wn = nweight(gi, nx, B.w1, B.w2, B.w3) * nweight(gj[p], ny, B.w1, B.w2, B.w3) *
                            nweight(k, nz, B.w1, B.w2, B.w3);
          if constexpr (IS_DBL) {
            const double mm = cabs_(np);
            acc_norm += wn * (mm * mm);
          } else {
            acc_norm += wn * (np.re * np.re + np.im * np.im);
          }
        } else {
          // Cash-Karp stages 2, 4, 6: kk = dt*f
          cx<T> kk;
          if constexpr (IS_DBL || !COMP) {
            kk = cmul(mk((T)B.dtr, (T)B.dti), f);
          } else {
            cx<double> kd = cmul(mk(B.dtr, B.dti), todbl(f));
            kk = mk((T)kd.re, (T)kd.im);
          }
          if constexpr (KB != K_CK_6) {
            stc<T>(B.kout, s, kk);
          } else {
            const cx<T> sc = COMP ? F[NF - 1] : oc[p][0];
            const cx<T> err = F[COMP ? 2 : 1] + rmul((T)ck_dc(6), kk);
            cx<double> np;
            if constexpr (IS_DBL) {
              np = F[0] + rmul((T)ck_c(6), kk);
              stc<T>(B.psi_out, s, np);
              const double e = cabs_(cdiv(err, sc));
              acc_emax = (e > acc_emax || e != e) ? e : acc_emax;
            } else {
              if constexpr (COMP) {
                const cx<double> dl = todbl(F[1]) + rmul(ck_c(6), todbl(kk));
                np = (todbl(F[0]) + todbl(oc[p][0])) + dl;
                const cx<T> hi = mk((T)np.re, (T)np.im);
                stc<T>(B.psi_out, s, hi);
                stc<T>(B.lo_out, s, mk((T)(np.re - (double)hi.re), (T)(np.im - (double)hi.im)));
                const double e2 = err_ratio2(err, sc);  // |err/scale|^2, as the one-stage kernel
                acc_emax = (e2 > acc_emax || e2 != e2) ? e2 : acc_emax;
              } else {
                const cx<T> nf = F[0] + rmul((T)ck_c(6), kk);
                stc<T>(B.psi_out, s, nf);
                np = todbl(nf);
                const float e = cabs_(cdiv(err, sc));  // abs(err/scale) (solve.f90:184)
                acc_emax = ((double)e > acc_emax || e != e) ? (double)e : acc_emax;
              }
            }
            const double wn = nweight(gi, nx, B.w1, B.w2, B.w3) * nweight(gj[p], ny, B.w1, B.w2, B.w3) *
                              nweight(k, nz, B.w1, B.w2, B.w3);
            if constexpr (IS_DBL) {
              const double mm = cabs_(np);
              acc_norm += wn * (mm * mm);
            } else {
              acc_norm += wn * (np.re * np.re + np.im * np.im);
            }
          }
        }
      }
    }

    // shift queues and FIFOs
#pragma unroll
    for (int m = 0; m < 2 * R; ++m)
#pragma unroll
      for (int p = 0; p < P; ++p) {
        qA[m][p] = qA[m + 1][p];
        qB[m][p] = qB[m + 1][p];
#pragma unroll
        for (int f = 0; f < NFA; ++f) FF[m][p][f] = FF[m + 1][p][f];
      }
#pragma unroll
    for (int m = 0; m < R; ++m) {
#pragma unroll
      for (int p = 0; p < P; ++p) wF[m][p] = wF[m + 1][p];
#pragma unroll
      for (int i = 0; i < HPT; ++i) G[m][i] = G[m + 1][i];
    }
  };
  // unrolled by D so that every prefetch buffer has a fixed register home
  for (int k = kb; k < z1; k += D) {
    step(k, rn[0], hn[0]);
    if constexpr (D > 1) {
      if (k + 1 < z1) step(k + 1, rn[1], hn[1]);
    }
    if constexpr (D > 2) {
      if (k + 2 < z1) step(k + 2, rn[2], hn[2]);
    }
    if constexpr (D > 3) {
      if (k + 3 < z1) step(k + 3, rn[3], hn[3]);
    }
  }
  static_assert(D >= 1 && D <= 4, "prefetch depth 1..4");

  // the compensated form maximised |err/scale|^2 (as the one-stage kernel)
  if constexpr (KB == K_CK_6 && COMP) acc_emax = sqrt(acc_emax);
  if constexpr (KindInfo<KB>::final_) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      acc_norm += __shfl_xor(acc_norm, o);
      double om = __shfl_xor(acc_emax, o);
      acc_emax = (om > acc_emax || om != om) ? om : acc_emax;
    }
    if (lane == 0) {
      red[0][wave] = acc_norm;
      red[1][wave] = acc_emax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double sn = 0.0, mx = 0.0;
      for (int w = 0; w < NW; ++w) {
        sn += red[0][w];
        mx = (red[1][w] > mx || red[1][w] != red[1][w]) ? red[1][w] : mx;
      }
      if (pn) pn[slot] = sn;
      if (pe) pe[slot] = mx;
    }
  }
}

// Single simulation: one workgroup per tile, XCD-aware order (as stage_kernel).
template <typename T, int R, int NW, int P, int D, int PAIR, bool COMP>
__global__ void __launch_bounds__(64 * NW) pair_kernel(const StageArgs A, const StageArgs B) {
  const int nb = B.nblocks;
  const int b = blockIdx.x;
  const int xcd = b & 7, loc = b >> 3;
  const int q8 = nb >> 3, r8 = nb & 7;
  const int L = xcd * q8 + (xcd < r8 ? xcd : r8) + loc;
  const int tx = L % B.ntx;
  const int rest = L / B.ntx;
  const int ty = rest % B.nty;
  const int tz = rest / B.nty;
  const int z0 = tz * B.zc;
  pair_tile<T, R, NW, P, D, PAIR, COMP>(A, B, tx, ty, z0, min(z0 + B.zc, B.nz), b, B.part_norm, B.part_emax);
}

// Many simulations: as stage_kernel_batched, with stage A's and stage B's
// descriptors of each instance.
template <typename T, int R, int NW, int P, int D, int PAIR, bool COMP>
__global__ void __launch_bounds__(64 * NW)
    pair_kernel_batched(const StageArgs *__restrict__ instA, const StageArgs *__restrict__ instB,
                        const WorkItem *__restrict__ work, int nwork, const int *__restrict__ active, double *pn,
                        double *pe) {
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const WorkItem it = work[w];
    if (active[it.inst])
      pair_tile<T, R, NW, P, D, PAIR, COMP>(instA[it.inst], instB[it.inst], it.tx, it.ty, it.z0, it.z1, it.slot,
                                            pn, pe);
    __syncthreads();
  }
}

}  // namespace gpe
