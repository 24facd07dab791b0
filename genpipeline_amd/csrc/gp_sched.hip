// Device-side scheduler: many independent problems, one launch per tick.
//
// Every slot carries a small state machine (SchedState) that walks the SIMP
// loop phase by phase: material law, Jacobi set-up, PCG iterations, sensitivity,
// filter, OC bisection, commit, compliance. One tick launch runs, for every
// slot, the phase its state names, on the slot's own block range, and block 0
// of the slot writes the state for the next tick. Slots therefore progress
// independently: a 32^3 problem does not wait for a 64^3 problem's PCG, and a
// slot that finishes can be refilled by the host between ticks.
//
// The state is double-buffered by tick parity: tick t reads st[t & 1] and
// writes st[(t + 1) & 1], so no block reads a state another block of the same
// launch is writing. Everything a phase needs from an earlier phase was
// written by an earlier launch.
//
// A slot's job uses as many blocks as its PCG sweep (nb_cg). Element and
// node phases spread over them; phases that leave per-block partial sums run
// the single-problem kernel's block count as virtual blocks on them, so every
// partial sum covers the same index set in the same order and a problem's
// results are bitwise those of Problem.run with the same K.u tile.
#include "gp_cg_body.hpp"
#include "gp_recycle.hpp"

namespace gp {

__device__ inline double sched_tol(const SchedState& s, const SchedCfg& c) {
  if (s.solve == SK_COMPL) return c.tol < 1e-6 ? c.tol : 1e-6;
  if (s.solve == SK_OC && c.tol_schedule)  // simp_solver_gpu.py:97-103
    return s.oc_it < 3 ? 1e-4 : (s.oc_it < 8 ? 1e-5 : 1e-6);
  return c.tol;
}

__device__ inline bool sched_use_x0(const SchedState& s, const SchedCfg& c) {
  return c.warm_start != 0 && (s.solve != SK_COMPL || c.warm_compl != 0);
}

// Last phase of a job: FP32 jobs copy their float iterate out, FP64-refined
// jobs already hold the caller's u.
__device__ inline int sched_end_phase(const SchedCfg& c) {
  return c.refine ? PH_DONE : PH_CAST_OUT;
}

// First phase after a finished solve.
__device__ inline int sched_post_solve(const SchedState& s) {
  return s.solve == SK_OC ? PH_SENS : PH_DOT;
}

// A warm solve of a job with recycled starts begins from the Galerkin
// combination once the ring holds a solution (gp_simp_run's order: the start
// replaces the last solution, then the solve proceeds as a warm one).
__device__ inline bool sched_recycled(const SchedState& s, const SchedCfg& c) {
  return c.recycle > 0 && s.rc_n > 0 && sched_use_x0(s, c);
}

// Phase after a solve ends, whatever its status: a recycling job stores the
// solution in its ring first.
__device__ inline int sched_solved(const SchedState& s, const SchedCfg& c) {
  return c.recycle > 0 ? PH_RC_STORE : sched_post_solve(s);
}

// Ring slot of stored solution i (0 = oldest) and the job's stored set.
__device__ inline int sched_ring_slot(const SchedState& s, const SchedCfg& c, int i) {
  return (s.rc_head - s.rc_n + i + c.recycle) % c.recycle;
}
__device__ inline RecycleSet sched_ring(const BatchProb& P, const SchedState& s,
                                        const SchedCfg& c) {
  RecycleSet rs;
  rs.n = s.rc_n;
  const long long nd = P.g.ndof;
  for (int i = 0; i < kRecycleMax; ++i) {
    rs.v[i] = i < s.rc_n ? P.rc_v + sched_ring_slot(s, c, i) * nd : nullptr;
    rs.kv[i] = i < s.rc_n ? P.rc_kv + i * nd : nullptr;
  }
  return rs;
}

// State after a finished OC update (iteration count oc_it already advanced).
__device__ inline void sched_after_update(SchedState& nx, const SchedCfg& c) {
  nx.phase = PH_STIFF;
  if (c.compl_each)
    nx.solve = SK_COMPL;
  else
    nx.solve = nx.oc_it < c.n_iters ? SK_OC : SK_FINAL;
}

// Next state of a slot after tick work; run by the slot's lead thread only.
// fin: the PCG solve (PH_CG) or the bisection (PH_OCSTEP) finished this tick.
// par: this tick's parity.
__device__ inline SchedState sched_next(const SchedState& cur, const SchedCfg& C,
                                        const BatchProb& P, bool fin, int par) {
  SchedState nx = cur;
  switch (cur.phase) {
    case PH_CAST_IN:
      nx.phase = PH_STIFF;
      break;
    case PH_STIFF:
      nx.phase = PH_DIAG;
      break;
    case PH_DIAG:
      if (C.refine) {
        // pcg_refine: a warm start keeps u (its fixed DOFs cleared in this
        // phase), otherwise u = 0; then the first residual
        nx.phase = sched_use_x0(cur, C) ? PH_RES : PH_Z64;
        nx.outer = 0;
        nx.ir_brk = 0;
      } else {
        nx.phase = sched_use_x0(cur, C) ? PH_KX0 : PH_CGINIT;
      }
      if (sched_recycled(cur, C)) {
        nx.phase = PH_RC_KV;
        nx.sub = 0;
      }
      break;
    case PH_RC_KV:
      if (cur.sub + 1 < cur.rc_n)
        nx.sub = cur.sub + 1;
      else
        nx.phase = PH_RC_GRAM;
      break;
    case PH_RC_GRAM:
      nx.phase = PH_RC_SOLVE;
      break;
    case PH_RC_SOLVE:
      nx.phase = PH_RC_COMB;
      break;
    case PH_RC_COMB:
      // the warm solve from the combined start
      nx.phase = C.refine ? PH_RES : PH_KX0;
      break;
    case PH_RC_STORE:
      nx.rc_head = (cur.rc_head + 1) % C.recycle;
      nx.rc_n = cur.rc_n + 1 < C.recycle ? cur.rc_n + 1 : C.recycle;
      nx.phase = sched_post_solve(cur);
      break;
    case PH_Z64:
      // outer < 0: f has no free part, the solve ends with u = 0
      nx.phase = cur.outer < 0 ? sched_solved(cur, C) : PH_RES;
      break;
    case PH_RES:
      nx.phase = PH_RESFIN;
      break;
    case PH_RESFIN: {
      // pcg_refine's decisions, in its order
      const double rr = P.sc->ir_rr, bb = P.sc->ir_bb;
      const double tol = sched_tol(cur, C);
      int st = -1;
      if (bb == 0.0) {
        nx.phase = PH_Z64;
        nx.outer = -1;
        break;
      }
      if (rr <= tol * tol * bb || !isfinite(rr))
        st = isfinite(rr) ? GP_OK : GP_ERR_NAN;
      else if (cur.ir_brk)
        st = GP_BREAKDOWN;
      else if (cur.outer >= C.max_outer)
        st = GP_NOT_CONVERGED;
      if (st >= 0) {
        if (st > nx.status) nx.status = st;
        nx.phase = sched_solved(cur, C);
      } else {
        nx.phase = PH_CGINIT;  // inner solve K d = r64 from d = 0
        nx.outer = cur.outer + 1;
      }
      break;
    }
    case PH_CORR:
      nx.phase = PH_RES;
      break;
    case PH_KX0:
      nx.phase = PH_CGINIT;
      break;
    case PH_CGINIT:
      nx.phase = PH_INITFIN;
      break;
    case PH_INITFIN:
      // launch k of a solve runs on a tick of parity k & 1: launch 0 waits
      // for an even tick
      nx.phase = par == 1 ? PH_CG : PH_ALIGN;
      nx.k = 0;
      break;
    case PH_ALIGN:
      nx.phase = PH_CG;
      break;
    case PH_CG:
      // the solve's last launch: the pending x and p updates first
      if (fin)
        nx.phase = PH_XPFIX;
      else
        nx.k = cur.k + 1;
      break;
    case PH_XPFIX:
      nx.pcg_total += P.sc->iter;
      if (C.refine) {
        // an inner solve: its correction is added whatever its status; a
        // breakdown ends the refinement after one more residual
        if (P.sc->status == GP_BREAKDOWN) nx.ir_brk = 1;
        nx.phase = PH_CORR;
      } else {
        if (P.sc->status > nx.status) nx.status = P.sc->status;
        nx.phase = sched_solved(cur, C);
      }
      break;
    case PH_SENS:
      nx.phase = PH_FILT;
      break;
    case PH_FILT:
      if (P.sc_oc->nan_flag) {
        // simp_solver_gpu.py:106-108: stop before the update
        nx.status = GP_ERR_NAN;
        nx.nan = 1;
        if (C.compl_each && cur.oc_it > 0) {
          nx.phase = sched_end_phase(C);
        } else {
          nx.phase = PH_STIFF;
          nx.solve = SK_FINAL;
        }
      } else {
        nx.phase = PH_OCMAX;
      }
      break;
    case PH_OCMAX:
      nx.phase = PH_OCINIT;
      break;
    case PH_OCINIT:
      nx.phase = PH_OCSTEP;
      nx.sub = 0;
      break;
    case PH_OCSTEP:
      if (fin) {
        nx.bis_total += P.sc_oc->oc_steps;
        nx.phase = PH_OCFINAL;
      } else {
        nx.sub = cur.sub + 1;
      }
      break;
    case PH_OCFINAL:
      nx.phase = PH_COMMIT;
      break;
    case PH_COMMIT:
      nx.oc_it = cur.oc_it + 1;
      sched_after_update(nx, C);
      break;
    case PH_DOT:
      nx.phase = PH_SUMFIN;
      break;
    case PH_SUMFIN: {
      const double c = P.sc_oc->sum;
      nx.curr_c = c;
      if (cur.solve == SK_COMPL) {
        // simp_solver_gpu.py:110-125
        bool stop = false;
        if (C.early_stop) {
          if (isfinite(cur.prev_c)) {
            const double ch = fabs(cur.prev_c - c) / (cur.prev_c + 1e-10);
            nx.small = ch < 0.01 ? cur.small + 1 : 0;
            stop = nx.small >= 3;
          }
          if (!stop) nx.prev_c = c;
        }
        if (stop || cur.oc_it >= C.n_iters) {
          nx.phase = sched_end_phase(C);
        } else {
          nx.phase = PH_STIFF;
          nx.solve = SK_OC;
        }
      } else {
        nx.phase = sched_end_phase(C);
      }
      break;
    }
    case PH_CAST_OUT:
      nx.phase = PH_DONE;
      break;
    default:
      break;
  }
  return nx;
}

// The recycled-start phases after the products K v (PH_RC_GRAM, PH_RC_SOLVE,
// PH_RC_COMB, PH_RC_STORE), on the job's blocks lb of nb. Gram and
// combination run the single-problem pass's gram_blocks layout as virtual
// blocks, so the start is bitwise gp_simp_run's. gl is the tick's LDS.
template <int N>
__device__ void sched_recycle_n(const BatchProb& P, const SchedCfg& C, const SchedState& cur,
                                int phase, int lb, int nb, GramLds& gl) {
  const RecycleSet rs = sched_ring(P, cur, C);
  const long long nd = P.g.ndof;
  const int gb = gram_blocks(nd);
  if (phase == PH_RC_GRAM) {
    for (int v = lb; v < gb; v += nb) body_gram<N>(rs, P.f, nd, P.rc_part, v, gb, gl);
  } else if (phase == PH_RC_SOLVE) {
    if (lb == 0) body_gram_solve<N>(gb, P.rc_part, P.rc_coef, gl);
  } else {
    // FP32 jobs start from the float combination, refined jobs from it
    // widened (gp_simp_run widens the float start, and pcg_refine clears the
    // supported DOFs of a warm start)
    body_combine<N>(rs, P.rc_coef, nd, C.refine ? nullptr : P.u, C.refine ? P.u64 : nullptr,
                    P.fixmask, lb, nb);
  }
}

// Kept out of line: the tick's sweep phases hold the kernel's register budget.
__device__ __attribute__((noinline)) void sched_recycle_phase(const BatchProb& P,
                                                              const SchedCfg& C,
                                                              const SchedState& cur, int phase,
                                                              int lb, int nb, GramLds& gl) {
  if (phase == PH_RC_STORE) {
    float* __restrict__ slot = P.rc_v + (long long)cur.rc_head * P.g.ndof;
    for (long long d = lb * (long long)kBlock + threadIdx.x; d < P.g.ndof;
         d += (long long)nb * kBlock)
      slot[d] = C.refine ? (float)P.u64[d] : P.u[d];
    return;
  }
  switch (cur.rc_n) {
    case 1: sched_recycle_n<1>(P, C, cur, phase, lb, nb, gl); break;
    case 2: sched_recycle_n<2>(P, C, cur, phase, lb, nb, gl); break;
    case 3: sched_recycle_n<3>(P, C, cur, phase, lb, nb, gl); break;
    case 4: sched_recycle_n<4>(P, C, cur, phase, lb, nb, gl); break;
    case 5: sched_recycle_n<5>(P, C, cur, phase, lb, nb, gl); break;
    case 6: sched_recycle_n<6>(P, C, cur, phase, lb, nb, gl); break;
    case 7: sched_recycle_n<7>(P, C, cur, phase, lb, nb, gl); break;
    case 8: sched_recycle_n<8>(P, C, cur, phase, lb, nb, gl); break;
    default: break;
  }
}

// Occupancy floor of 4 waves per SIMD: the 128-register budget the PCG sweep
// fits in, kept for every phase of the kernel (the FP64 sensitivity phase
// spills a little; a fourth wave per SIMD is worth more to the sweep).
// PAR: the tick's parity (t & 1). The PCG phase updates x and p two
// iterations at a time, as the single-problem solve does: launch k of a solve
// runs on a tick of parity k & 1 (PH_ALIGN), so odd ticks leave x and p (XP 1)
// and even ticks apply both pending updates (XP 2), one sweep instantiation
// per kernel. Launch 0 runs XP 2 with alpha_{-2} = beta_{-2} = 0 (the body
// reads them only from launch 1 on) and r_{-2} = 0 (cleared in PH_CGINIT): its
// x and p are the every-iteration form's values.
template <int TJ, int TK, int PAR>
__global__ __launch_bounds__(kBlock, 4) void k_tick(SchedTable tab, int t) {
  static_assert(TJ * TK <= kBlock, "the scheduler runs every body with kBlock threads");
  static_assert(sizeof(CornerLds<float, TJ * TK>) >= sizeof(GramLds),
                "the recycled-start phases borrow the corner buffer");
  __shared__ __attribute__((aligned(16))) CornerLds<float, TJ * TK> cv;
  const int xb = xcd_block(blockIdx.x, gridDim.x);
  const int gb = tab.act ? tab.act[xb] : xb;
  const int b = tab.b2s[gb];
  SchedSlot& S = tab.slots[b];
  const int lb = gb - tab.off[b];
  const SchedState& cur = S.st[t & 1];
  const BatchProb& P = S.P;
  const int phase = cur.phase;
  const int nb = S.nb_cg;  // blocks of this job; the rest of the range idles
  bool fin = false;

  switch (lb < nb ? phase : PH_IDLE) {
    case PH_CAST_IN:
      for (long long d = lb * (long long)kBlock + threadIdx.x; d < P.g.ndof;
           d += (long long)nb * kBlock)
        P.u[d] = (float)P.u64[d];
      break;
    case PH_STIFF:
      body_stiffness(P.g.ne, P.xPhys, P.E, P.law, P.penal, P.E0, P.Emin, lb, nb);
      break;
    case PH_DIAG:
      body_diag<float>(P.g, P.ke32, P.E, P.fixmask, P.minv, 2, lb, nb);  // per node
      if (S.cfg.refine && sched_use_x0(cur, S.cfg))
        for (int n = lb * kBlock + threadIdx.x; n < P.g.nn; n += nb * kBlock) {
          const uint8_t fm = P.fixmask[n];
          if (fm & 1) P.u64[3ll * n] = 0.0;
          if (fm & 2) P.u64[3ll * n + 1] = 0.0;
          if (fm & 4) P.u64[3ll * n + 2] = 0.0;
        }
      break;
    case PH_Z64:
      for (long long d = lb * (long long)kBlock + threadIdx.x; d < P.g.ndof;
           d += (long long)nb * kBlock)
        P.u64[d] = 0.0;
      break;
    case PH_RES:
      for (int v = lb; v < S.nb_vec; v += nb)
        body_res64(P.g, P.ke64, P.E, P.u64, P.f, P.fixmask, P.r64, P.parts.A, P.parts.RZ[1], v,
                   S.nb_vec);
      break;
    case PH_RESFIN:
      if (lb == 0) body_res_fin(P.parts, S.nb_vec, P.sc);
      break;
    case PH_CORR:
      for (long long d = lb * (long long)kBlock + threadIdx.x; d < P.g.ndof;
           d += (long long)nb * kBlock)
        P.u64[d] += (double)P.u[d];
      break;
    case PH_KX0:
    case PH_RC_KV: {
      // w0 = K u of a warm start, or K v of stored solution sub (one sweep
      // site for both)
      const float* src = P.u;
      float* dst = P.wb[0];
      if (phase == PH_RC_KV) {
        src = P.rc_v + (long long)sched_ring_slot(cur, S.cfg, cur.sub) * P.g.ndof;
        dst = P.rc_kv + (long long)cur.sub * P.g.ndof;
      }
      body_ku_sweep<float, TJ, TK, kBlock>(P.g, P.ni, P.ntj, P.ntk, P.mc, P.E, src, dst,
                                           P.fixmask, lb, cv);
      break;
    }
    case PH_RC_GRAM:
    case PH_RC_SOLVE:
    case PH_RC_COMB:
    case PH_RC_STORE:
      sched_recycle_phase(P, S.cfg, cur, phase, lb, nb, *reinterpret_cast<GramLds*>(&cv));
      break;
    case PH_CGINIT: {
      // refined jobs: the inner solve of K d = r64 from d = 0
      const bool ref = S.cfg.refine != 0;
      const int x0 = !ref && sched_use_x0(cur, S.cfg) ? 1 : 0;
      for (int v = lb; v < S.nb_vec; v += nb)
        body_cg_init<float>(P.g.nn, ref ? P.r64 : P.f, P.wb[0], x0, P.u, P.rb[1], P.sb[0],
                            P.wb[1], P.pc, P.minv, P.fixmask, 1, P.parts, v, S.nb_vec);
      // r_{-2} for launch 0's XP 2 form: finite (zero) whatever a former
      // solve left in this buffer
      for (long long d = lb * (long long)kBlock + threadIdx.x; d < P.g.ndof;
           d += (long long)nb * kBlock)
        P.rb[0][d] = 0.0f;
      if (lb == 0 && threadIdx.x == 0) P.sc_oc->nan_flag = 0;
      break;
    }
    case PH_INITFIN:
      if (lb == 0)
        body_pcg_init_fin(P.parts, S.nb_vec,
                          S.cfg.refine ? S.cfg.inner_tol : sched_tol(cur, S.cfg), S.cfg.max_iter,
                          P.sc);
      break;
    case PH_CG: {
      const int k = cur.k, c = k & 1, o = c ^ 1;  // k & 1 == PAR (PH_ALIGN)
      fin = body_cg_modal<float, TJ, TK, float, kBlock, false, PAR == 0 ? 2 : 1>(
          P.g, P.ni, P.ntj, P.ntk, P.mc, P.E, P.rb[o], P.wb[o], P.sb[c], P.minv, P.rb[c], P.wb[c],
          P.sb[o], P.u, P.pc, P.fixmask, P.sc, P.parts, k, lb, nb, cv);
      break;
    }
    case PH_XPFIX:
      body_cg_xp_fix<float>(P.g.nn, P.sc, P.rb[0], P.minv, P.u, P.pc, lb, nb);
      break;
    case PH_SENS:
      if (S.cfg.refine)
        body_sens<double, true>(P.g, P.ke64, P.mc64, P.xPhys, P.u64, P.law, P.penal, P.E0,
                                P.Emin, P.dc, 0, nullptr, P.sc_oc, lb, nb);
      else
        body_sens<float, true>(P.g, P.ke64, P.mc64, P.xPhys, P.u, P.law, P.penal, P.E0, P.Emin,
                               P.dc, 0, nullptr, P.sc_oc, lb, nb);
      break;
    case PH_FILT:
      body_filter(P.g, P.st, P.Hs, P.dc, P.dcf, 1, lb, nb);
      break;
    case PH_OCMAX:
      if (P.oc_variant == GP_OC_GPU)
        for (int v = lb; v < S.nb_oc; v += nb) body_oc_max(P.g.ne, P.dcf, P.parts.M, v, S.nb_oc);
      break;
    case PH_OCINIT:
      if (lb == 0) body_oc_init_fin(P.parts, S.nb_oc, P.oc_variant, P.volfrac, P.move, P.sc_oc);
      break;
    case PH_OCSTEP:
      // every virtual block reaches the same decision; the lead's is used
      for (int v = lb; v < S.nb_oc; v += nb) {
        const bool f = body_oc_step(cur.sub, P.g.ne, P.x, P.dcf, P.w, P.pres, P.n_pres, P.parts,
                                    P.sc_oc, v, S.nb_oc);
        if (v == lb) fin = f;
      }
      break;
    case PH_OCFINAL:
      body_oc_final(P.g.ne, P.x, P.dcf, P.pres, P.sc_oc, P.xnew, lb, nb);
      break;
    case PH_COMMIT:
      // every element is filtered, pinned and committed by the one thread
      // that owns it; the stencil reads only xnew, complete since last tick
      body_filter(P.g, P.st, P.Hs, P.xnew, P.xphys_new, 0, lb, nb);
      if (P.pres != nullptr) body_set_preserved(P.g.ne, P.pres, P.xphys_new, lb, nb);
      for (int e = lb * kBlock + threadIdx.x; e < P.g.ne; e += nb * kBlock) {
        P.x[e] = P.xnew[e];
        P.xPhys[e] = P.xphys_new[e];
      }
      break;
    case PH_DOT:
      for (int v = lb; v < S.nb_vec; v += nb)
        if (S.cfg.refine)
          body_dot<double, double>(P.g.ndof, P.f, P.u64, P.parts.S, v, S.nb_vec);
        else
          body_dot<double, float>(P.g.ndof, P.f, P.u, P.parts.S, v, S.nb_vec);
      break;
    case PH_SUMFIN:
      if (lb == 0) body_sum_fin(P.parts.S, S.nb_vec, &P.sc_oc->sum);
      break;
    case PH_CAST_OUT:
      for (long long d = lb * (long long)kBlock + threadIdx.x; d < P.g.ndof;
           d += (long long)nb * kBlock)
        P.u64[d] = (double)P.u[d];
      break;
    default:
      break;
  }
  if (lb == 0 && threadIdx.x == 0) {
    const SchedState nx = sched_next(S.st[t & 1], S.cfg, P, fin, PAR);
    S.st[(t + 1) & 1] = nx;
    tab.status[b] = nx;
  }
}

// Tiles of up to kBlock lanes; a tile of fewer lanes (14 x 18 = 252, 7 x 34 =
// 238, 6 x 34 = 204) runs in kBlock-thread blocks whose extra lanes idle in
// the sweeps.
bool sched_tick(const SchedTable& tab, int tj, int tk, int blocks, int t, hipStream_t st) {
#define GP_TILE(TJ_, TK_)                                                                   \
  if (tj == TJ_ && tk == TK_) {                                                             \
    if (t & 1)                                                                              \
      hipLaunchKernelGGL((k_tick<TJ_, TK_, 1>), dim3(blocks), dim3(kBlock), 0, st, tab, t); \
    else                                                                                    \
      hipLaunchKernelGGL((k_tick<TJ_, TK_, 0>), dim3(blocks), dim3(kBlock), 0, st, tab, t); \
    return true;                                                                            \
  }
  GP_TILE(8, 32)
  GP_TILE(16, 16)
  GP_TILE(14, 18)
  GP_TILE(6, 34)
  GP_TILE(7, 34)
#undef GP_TILE
  return false;
}

}  // namespace gp
