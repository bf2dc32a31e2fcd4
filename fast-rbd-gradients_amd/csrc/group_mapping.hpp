// Knot groups of L lanes (L = 4, 7 or 8 lanes per knot): the mappings between
// one lane per knot (a knot's whole chain in one lane, 64 knots per wave) and
// the 16-lane groups (4 knots per wave).  A wave holds floor(64 / L) knots;
// with L = 7 that is 9 knots on 63 lanes, and each lane has exactly one joint's
// sin/cos, one column of Minv and two of the 14 column tasks (7 d/dqd columns,
// 6 d/dq columns, qdd) -- no lane idles in any stage.  So a batch of ~9,000
// knots is one wave per SIMD of the device (1,024 on MI355X) instead of two
// rounds of 16-lane groups, and each lane's chain is the shared steps plus two
// columns instead of thirteen.
//
// Per lane (role r of its knot's group):
//   1. sin/cos of the joints r, r + L, ... (< 7), the other joints' values read
//      from the group's lanes by shuffles (knot_trig_group for any L);
//   2. Minv: with L >= 7 lane r computes column r (compute_minv_column) into
//      the group's LDS block and every lane reads the whole matrix back; with
//      L < 7 every lane runs the whole recursion (compute_minv), which costs
//      less than several column recursions;
//   3. the bias RNEA, qdd = Minv (u - c) and the RNEA with qdd (shared steps,
//      phase_a_after_minv);
//   4. tasks r, r + L, ... : each a derivative column (dtau_column_uniform,
//      the same instructions in every lane whatever its column) times -Minv,
//      stored, or (task 13) qdd and the zero first d/dq column.
// The arithmetic is the 16-lane groups' (the same functions), so the results
// agree with them to FP32 rounding; for a fixed L they do not depend on which
// other knots share the wave.  Plain C++ when compiled for the host, where the
// test harness runs the stages of every lane role one after another.
#pragma once
#include "fdgrad_knot.hpp"

namespace rbd {

// The L-lane group kernels run in workgroups of 1, 2 or 4 independent waves
// (a launch parameter); their LDS is sized for this many.
constexpr int kMaxGroupWaves = 4;

template <int L>
struct Group {
  static_assert(L >= 2 && L <= 16, "lanes per knot");
  static constexpr int kSlots = 64 / L;                   // knots per wave
  static constexpr int kLanesUsed = kSlots * L;            // 63 for L = 7
  static constexpr int kTrigRounds = (NQ + L - 1) / L;     // joints per lane
  static constexpr int kTasks = 2 * NQ;                    // 7 + 6 columns and qdd
  static constexpr int kRounds = (kTasks + L - 1) / L;     // tasks per lane
  static constexpr bool kMinvColumns = L >= NQ;            // one Minv column per lane (else whole per lane)
};

// A lane's knot slot in the wave and role in the knot's group.  Lanes past
// kLanesUsed (lane 63 for L = 7) join the last slot as copies of its role 0:
// they compute the same values as that lane and store nothing.
template <int L>
RBD_HD constexpr int group_slot(int lane) {
  return lane / L < Group<L>::kSlots ? lane / L : Group<L>::kSlots - 1;
}
template <int L>
RBD_HD constexpr int group_role(int lane) {
  return lane < Group<L>::kLanesUsed ? lane % L : 0;
}

// Stage 1: sin/cos of the knot's joints (lane role r computes joints r, r + L,
// ...; the group's lanes read the rest from each other by shuffles; base is
// the group's first lane) and the lane's share of Minv: its column into the
// group's LDS block mi (L >= 7, roles < 7; written only when `store`), or the
// whole matrix into Mi (L < 7).
template <class P, int L, class In>
RBD_HD RBD_INLINE void group_stage1(const In& in, int base, int role, bool store, CS<KnotTrig<P>> tw[NQ],
                                    typename P::Minv* mi, typename P::Minv Mi[NQ][NQ]) {
  using SM = typename P::Minv; using SP = typename P::Pivot;
#if defined(__HIP_DEVICE_COMPILE__)
  using ST = KnotTrig<P>;
  ST s[Group<L>::kTrigRounds], c[Group<L>::kTrigRounds];
  static_for<0, Group<L>::kTrigRounds>([&](auto Tc) {
    constexpr int t = decltype(Tc)::value;
    const int j = role + t * L;
    knot_sincos<P>(ST(in.q(j < NQ ? j : 0)), &s[t], &c[t]);
  });
  static_for<0, NQ>([&](auto Kc) {
    constexpr int k = decltype(Kc)::value;
    tw[k] = CS<ST>{__shfl(c[k / L], base + k % L, 64), __shfl(s[k / L], base + k % L, 64)};
  });
#else
  (void)base;
  knot_trig<P>(in, tw);
#endif
  CS<SM> tm[NQ];
  for (int k = 0; k < NQ; ++k) tm[k] = CS<SM>{SM(tw[k].c), SM(tw[k].s)};
  if constexpr (Group<L>::kMinvColumns) {
    const int j = role < NQ ? role : NQ - 1;
    SM col[NQ];
    compute_minv_column<SM, SP>(tm, j, col);
    if (store && role < NQ)
      for (int k = 0; k < NQ; ++k)
        if (k <= j) mi[upper_index(k, j)] = col[k];
    (void)Mi;
  } else {
    (void)mi; (void)store; (void)role;
    compute_minv<SM, SP>(tm, Mi);
  }
}

// Stage 2 (after every lane's stage 1; with L >= 7 Mi is read from mi here):
// the shared steps, then the lane's tasks.
template <class P, int L, class In, class Out>
RBD_HD RBD_INLINE void group_stage2(const In& in, Out& out, const Options& opt, int role, const CS<KnotTrig<P>> tw[NQ],
                                    const typename P::Minv* mi, typename P::Minv Mi[NQ][NQ]) {
  using SC = typename P::Col;
  if constexpr (Group<L>::kMinvColumns) load_packed_minv(mi, Mi);
  PhaseC<SC> PC;
  typename P::Rnea qdd[NQ];
  phase_a_after_minv<P>(in, opt, tw, Mi, PC, qdd);
  static_for<0, Group<L>::kRounds>([&](auto Tc) {
    constexpr int t = decltype(Tc)::value;
    const int task = role + t * L;                 // 0..6 d/dqd c; 7..12 d/dq c - 6; 13 qdd; more: nothing
    if (task == kRoleQdd) {
      for (int r = 0; r < NQ; ++r) out.qdd(r, float(qdd[r]));
      for (int r = 0; r < NQ; ++r) out.dq(r, 0, 0.0f);
    }
    const bool dq = task >= NQ;
    const int c = dq ? task - (NQ - 1) : task;     // past the last column: computed, not stored
    SC tau[NQ], col[NQ];
    dtau_column_uniform(PC, c, dq, tau);
    add_on_diagonal(tau, dq ? -1 : c, SC(opt.damping));
    minv_times<true>(PC.Mi, tau, col);
    if (task < NQ) {
      for (int r = 0; r < NQ; ++r) out.dqd(r, c, float(col[r]));
    } else if (task < 2 * NQ - 1) {
      for (int r = 0; r < NQ; ++r) out.dq(r, c, float(col[r]));
    }
  });
}

// The full path of one knot on its group of L lanes (device code): lane
// `lane` of the wave, role group_role<L>(lane); mi is the group's LDS block
// (kMinvPacked values, used when L >= 7); out is masked for lanes that store
// nothing.  The caller's next use of mi must come after a wave_lds_sync.
template <class P, int L, class In, class Out>
RBD_HD RBD_INLINE void fdgrad_knot_group(const In& in, Out& out, const Options& opt, int lane, typename P::Minv* mi) {
  const int role = group_role<L>(lane);
  const int base = group_slot<L>(lane) * L;
  CS<KnotTrig<P>> tw[NQ];
  typename P::Minv Mi[NQ][NQ];
  group_stage1<P, L>(in, base, role, lane < Group<L>::kLanesUsed, tw, mi, Mi);
  if constexpr (Group<L>::kMinvColumns) wave_lds_sync();
  group_stage2<P, L>(in, out, opt, role, tw, mi, Mi);
}

}  // namespace rbd
