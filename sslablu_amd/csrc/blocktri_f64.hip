// Block-tridiagonal S-system factor / solve on gfx950 (FP64).
//
//   T_i x_i + SiM_i x_{i-1} + SiP_i x_{i+1} = v_i ,  i = 0..nb-1   (non-cyclic)
//
// Red-black (cyclic) reduction, restated for a GPU:  at every level the odd
// nodes are eliminated and the even nodes form the next, half-size level, as in
// direct_solve/omsdirectsolve.py:283-333 (RedBlackSolver.build_block_RB_solver_level),
// generalised to any block count (an even node at the right end simply has no
// right neighbour) so no identity padding is needed.  Each level keeps only what
// the solve needs, with explicit inverses so every solve step is a GEMM:
//
//   odd  j = 2c+1 :  TG_c = [ Tinv_j | Tinv_j SiM_j | Tinv_j SiP_j ]      (m x 3m)
//   even i = 2b   :  WW_b = [ SiM_i | SiP_i ]                              (m x 2m)
//   next level    :  T'_b   = T_i - SiM_i (Tinv SiP)_{i-1} - SiP_i (Tinv SiM)_{i+1}
//                    SiM'_b = -SiM_i (Tinv SiM)_{i-1},   SiP'_b = -SiP_i (Tinv SiP)_{i+1}
//
// Every product of the next level reuses TG, so a level costs the inverses, TG's 4 m^3
// per odd node and 8 m^3 per even node; forming the reference's W = SiM_i Tinv_{i-1} as
// well (for a one-launch forward step) would add another 4 m^3 per even node, a fifth
// of the factor.
//
// Tinv_j (and the coarsest block's inverse) come from inv_batched: block elimination on
// MFMA GEMMs for all odd blocks of a level at once, with a sampled residual check that
// falls back to the pivoted LU (the reference's lu_factor / lu_solve(., I)).
//
// Solve, per level, forward (two launches):  y_j = Tinv_j v_j for odd j, then
// v_i -= SiM_i y_{i-1} + SiP_i y_{i+1} for even i;  backward (one launch):
// x_j = y_j - (Tinv SiM)_j x_{j-1} - (Tinv SiP)_j x_{j+1}.  Level-l node k lives
// in global block slot k << l, so every level works in place on V and X.
//
// The reduction stops (after at least one level) at a coarsest level of
// n_top <= min(4, 2048 / m) nodes, whose (n_top m)^2 system is stored as an explicit
// inverse: the solve's deepest levels are a few small, latency-bound launches each, and
// one grouped product over the dense inverse replaces all of them.  The inverse itself
// comes from a red-black factorization of that small system down to one block, solved
// for the identity (n_top m right-hand sides: large, efficient products), not from one
// LU of the whole (n_top m)^2 matrix, whose panels would run on a single workgroup.
//
// Thomas (direct_solve/omsdirectsolve.py:48-141) is kept as the faithful
// sequential variant: A_i = SiM_i B_{i-1}^{-1}, B_i = T_i - A_i SiP_{i-1}.
#include "blocktri_f64.hpp"

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "gemm.hpp"
#include "inv.hpp"
#include "lu_f64.hpp"

namespace sslab {

namespace {

struct RbLevel {
  int n, nE, nO;
  long long tg_off;  // factor offset (doubles) of TG blocks (nO x m x 3m)
  long long ww_off;  // factor offset of WW blocks (nE x m x 2m)
};

struct RbPlan {
  int nb, m, L;
  std::vector<RbLevel> lv;
  int n_top;               // nodes of the coarsest level
  long long top_off;       // inverse of the coarsest level's system ((n_top m)^2)
  long long factor_doubles;
  // workspace (doubles): two ping-pong level sets (T', SiM', SiP'), LU blocks,
  // the stacked [I | SiM | SiP] right-hand sides.
  long long set_blocks;    // blocks per array in one ping-pong set
  long long work_doubles;
  long long work_ints;     // ipiv + perm + info, then the deferred inverse checks' flags
  long long flag_off;      // int offset of those flags (one per inverted block, as info)
  int max_odd;
  int piv_blocks;          // ipiv / perm capacity in blocks of m: max(max_odd, n_top)
};

// Largest coarsest level inverted as one dense system: at most 4 nodes and 2048 rows.
inline int rb_top_max(int m) {
  const int k = 2048 / m;
  return k < 1 ? 1 : (k > 4 ? 4 : k);
}

// top_max: the largest coarsest level kept as a dense inverse (0: rb_top_max(m)).
RbPlan rb_plan(int nb, int m, int top_max = 0) {
  RbPlan p{};
  p.nb = nb;
  p.m = m;
  const long long m2 = (long long)m * m;
  long long off = 0;
  int n = nb;
  p.max_odd = 1;
  if (top_max <= 0) top_max = rb_top_max(m);
  while (n > 1 && (p.lv.empty() || n > top_max)) {
    RbLevel l{};
    l.n = n;
    l.nE = (n + 1) / 2;
    l.nO = n / 2;
    l.tg_off = off;
    off += (long long)l.nO * 3 * m2;
    l.ww_off = off;
    off += (long long)l.nE * 2 * m2;
    if (l.nO > p.max_odd) p.max_odd = l.nO;
    p.lv.push_back(l);
    n = l.nE;
  }
  p.L = (int)p.lv.size();
  p.n_top = n;
  p.top_off = off;
  const long long nt2 = (long long)n * n * m2;
  off += nt2;
  p.factor_doubles = off;
  p.set_blocks = (nb + 1) / 2;
  // ping-pong sets, LU blocks, [I | SiM | SiP] right-hand sides, then (n > 1) the
  // coarsest level's own red-black factor and workspace and its identity right-hand side
  p.work_doubles = 2 * 3 * p.set_blocks * m2 + (long long)p.max_odd * m2 +
                   (long long)p.max_odd * 3 * m2;
  p.piv_blocks = p.max_odd > n ? p.max_odd : n;
  p.work_ints = (long long)p.piv_blocks * m * 2 + p.piv_blocks;
  if (n > 1) {
    const RbPlan sp = rb_plan(n, m, 1);
    p.work_doubles += sp.factor_doubles + sp.work_doubles + nt2;
    if (sp.work_ints > p.work_ints) p.work_ints = sp.work_ints;
  }
  p.flag_off = p.work_ints;
  p.work_ints += rb_info_count(nb);
  return p;
}

__global__ void k_fill_int(int* __restrict__ p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void zero_int_async(int* p, int n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3(ceil_div(n, 256)), dim3(256), 0, st, p, n, 0);
}

}  // namespace

int rb_shape(int nb, int m, int* nodd, int nodd_cap, int* n_top) {
  if (nb < 1 || m < 1) return -1;
  const RbPlan p = rb_plan(nb, m);
  if (nodd)
    for (int l = 0; l < p.L && l < nodd_cap; ++l) nodd[l] = p.lv[l].nO;
  if (n_top) *n_top = p.n_top;
  return p.L;
}

int rb_sizes(int nb, int m, long long* factor_doubles, long long* work_doubles,
             long long* work_ints) {
  SSLAB_REQUIRE(nb >= 1 && m >= 1, "rb_sizes: nb and m must be positive");
  RbPlan p = rb_plan(nb, m);
  if (factor_doubles) *factor_doubles = p.factor_doubles;
  if (work_doubles) *work_doubles = p.work_doubles;
  if (work_ints) *work_ints = p.work_ints;
  return kOk;
}

// Whole red-black solve of a system planned with the given coarsest-level limit (defined
// with the solve sweeps below; the factorization uses it for the coarsest level's inverse).
template <typename T>
static int rb_solve_sub(int nb, int m, int r, const T* factor, T* V, int ldv, T* X, int ldx,
                        hipStream_t st, int top_max = 1);

// inv_mode: kInvDeferred (every inverse's check writes its flags at `flags`, one int per
// inverted block in the layout of info, for the caller to read) or kInvLU (pivoted LU
// inverses only).
template <typename E>
static int rb_factor_t(int nb, int m, const E* T, const E* SiM, const E* SiP, E* factor, E* work,
                       int* iwork, int* info, hipStream_t st, int top_max, int inv_mode,
                       int* flags) {
  using GemmLaunch = GemmLaunchT<E>;
  using GemmDesc = GemmDescT<E>;
  SSLAB_REQUIRE(nb >= 1 && m >= 1, "rb_factor: nb and m must be positive");
  SSLAB_REQUIRE(factor && work && iwork, "rb_factor: null pointer");
  // T == NULL: identity diagonal blocks (the oms S-system u_i + S u = rhs).  Level 0 then
  // needs no inverse: T^{-1} = I, TG = [I | SiM | SiP], T' = I - SiM SiP - SiP SiM.
  SSLAB_REQUIRE(nb == 1 || (SiM && SiP), "rb_factor: null off-diagonal blocks");
  const RbPlan p = rb_plan(nb, m, top_max);
  const long long m2 = (long long)m * m;
  const long long setsz = 3 * p.set_blocks * m2;
  E* sets[2] = {work, work + setsz};
  // the inverses' workspace (4 m^2 per odd node: recursion, residual sample, or the LU
  // fallback's factors and identity), then the coarsest level's
  E* LUw = work + 2 * setsz;
  E* Dw = LUw + (long long)p.max_odd * 4 * m2;
  int* ipiv = iwork;  // the inverses' ints: pivots, permutations, check flags
  // info layout: one int per odd node per level, then one for the top block.

  const E* cT = T;
  const E* cM = SiM;
  const E* cP = SiP;
  int info_base = 0;
  for (int l = 0; l < p.L; ++l) {
    const RbLevel& v = p.lv[l];
    const int n = v.n, nE = v.nE, nO = v.nO;
    E* TG = factor + v.tg_off;
    E* WW = factor + v.ww_off;
    E* nT = sets[(l + 1) & 1];
    E* nM = nT + p.set_blocks * m2;
    E* nP = nM + p.set_blocks * m2;

    const bool ident = (l == 0 && T == nullptr);
    const int nOr = (n % 2 == 0) ? nO - 1 : nO;  // odd nodes with a right neighbour
    if (ident) {
      // TG_j = [I | SiM_j | SiP_j],  W_b = [SiM_{2b} | SiP_{2b}],  T'_b = I (then updated).
      SSLAB_TRY(set_identity_batched(m, TG, 3 * m, 3 * m2, nO, st));
      SSLAB_TRY(copy_block_batched(m, m, cM + m2, m, 2 * m2, TG + m, 3 * m, 3 * m2, nO, st));
      SSLAB_TRY(copy_block_batched(m, m, cP + m2, m, 2 * m2, TG + 2 * m, 3 * m, 3 * m2, nOr, st));
      if (nOr < nO)
        SSLAB_TRY(zero_block_batched(m, m, TG + (long long)nOr * 3 * m2 + 2 * m, 3 * m, 3 * m2, 1,
                                     st));
      if (info) zero_int_async(info + info_base, nO, st);
      info_base += nO;
      SSLAB_TRY(set_identity_batched(m, nT, m, m2, nE, st));
    } else {
      // 1-3. Tinv_j (first block of TG): batched block-elimination inverse of the odd
      //      diagonal blocks (pivoted LU when its sampled residual fails), then
      // 4.   TG_j[:, m:3m] = Tinv_j [SiM_j | SiP_j] as one MFMA GEMM launch (SiP of a
      //      last odd node without a right neighbour is zero).
      SSLAB_TRY(inv_batched<E>(m, cT + m2, m, 2 * m2, TG, 3 * m, 3 * m2, nO, LUw, ipiv,
                               info ? info + info_base : nullptr, 0.0, nullptr, st, inv_mode,
                               flags ? flags + info_base : nullptr));
      info_base += nO;
      {
        GemmLaunch G{};
        G.nd = 0;
        G.d[G.nd] = gemm_desc(m, m, nO, TG + m, 3 * m, 3 * m2, nullptr, 0, 0, 0.0);
        gemm_add_term(G.d[G.nd++], gemm_term(TG, 3 * m, 3 * m2, 0, cM + m2, m, 2 * m2, 0, m, 1.0));
        if (nOr > 0) {
          G.d[G.nd] = gemm_desc(m, m, nOr, TG + 2 * m, 3 * m, 3 * m2, nullptr, 0, 0, 0.0);
          gemm_add_term(G.d[G.nd++],
                        gemm_term(TG, 3 * m, 3 * m2, 0, cP + m2, m, 2 * m2, 0, m, 1.0));
        }
        SSLAB_TRY(gemm_launch(G, st));
      }
      if (nOr < nO)
        SSLAB_TRY(zero_block_batched(m, m, TG + (long long)nOr * 3 * m2 + 2 * m, 3 * m, 3 * m2, 1,
                                     st));
    }
    // 5. WW_b = [SiM_{2b} | SiP_{2b}]: the even nodes' own couplings, for the forward
    //    sweep (SiM_0 and, for odd n, the last even node's SiP do not exist).
    if (nE > 1)
      SSLAB_TRY(copy_block_batched(m, m, cM + 2 * m2, m, 2 * m2, WW + 2 * m2, 2 * m, 2 * m2,
                                   nE - 1, st));
    SSLAB_TRY(copy_block_batched(m, m, cP, m, 2 * m2, WW + m, 2 * m, 2 * m2, nO, st));
    // 6. Next level: T', SiM', SiP' from the even nodes' couplings and TG of their odd
    //    neighbours (left of even node b: odd node b - 1; right: odd node b).
    {
      const E* TGM = TG + m;      // (Tinv SiM)_c, stride 3 m^2
      const E* TGP = TG + 2 * m;  // (Tinv SiP)_c
      const long long s3 = 3 * m2;
      GemmLaunch G{};
      int k = 0;
      const int both = (nE < nO ? nE : nO) - 1;  // b in [1, min(nE,nO))
      // Cin: T_{2b} (strided by two blocks), or the identity just written into nT.
      const E* cin = ident ? nT : cT;
      const long long scin = ident ? m2 : 2 * m2;
      if (both > 0) {
        GemmDesc d = gemm_desc(m, m, both, nT + m2, m, m2, cin + scin, m, scin, 1.0);
        gemm_add_term(d, gemm_term(cM + 2 * m2, m, 2 * m2, 0, TGP, 3 * m, s3, 0, m, -1.0));
        gemm_add_term(d, gemm_term(cP + 2 * m2, m, 2 * m2, 0, TGM + s3, 3 * m, s3, 0, m, -1.0));
        G.d[k++] = d;
      }
      {  // b = 0: right neighbour only (n >= 2 so node 1 exists)
        GemmDesc d = gemm_desc(m, m, 1, nT, m, m2, cin, m, m2, 1.0);
        gemm_add_term(d, gemm_term(cP, m, m2, 0, TGM, 3 * m, s3, 0, m, -1.0));
        G.d[k++] = d;
      }
      if (nE > nO && nE - 1 >= 1) {  // n odd: last even node has a left neighbour only
        const long long b = nE - 1;
        GemmDesc d = gemm_desc(m, m, 1, nT + b * m2, m, m2, cin + b * scin, m, m2, 1.0);
        gemm_add_term(d, gemm_term(cM + 2 * b * m2, m, m2, 0, TGP + (b - 1) * s3, 3 * m, s3, 0, m,
                                   -1.0));
        G.d[k++] = d;
      }
      if (nE - 1 > 0) {  // SiM'_b = -SiM_{2b} (Tinv SiM)_{2b-1}, b in [1, nE)
        GemmDesc d = gemm_desc(m, m, nE - 1, nM + m2, m, m2, nullptr, 0, 0, 0.0);
        gemm_add_term(d, gemm_term(cM + 2 * m2, m, 2 * m2, 0, TGM, 3 * m, s3, 0, m, -1.0));
        G.d[k++] = d;
      }
      if (nE - 1 > 0) {  // SiP'_b = -SiP_{2b} (Tinv SiP)_{2b+1}, b in [0, nE-1)
        GemmDesc d = gemm_desc(m, m, nE - 1, nP, m, m2, nullptr, 0, 0, 0.0);
        gemm_add_term(d, gemm_term(cP, m, 2 * m2, 0, TGP, 3 * m, s3, 0, m, -1.0));
        G.d[k++] = d;
      }
      G.nd = k;
      SSLAB_TRY(gemm_launch(G, st));
    }
    cT = nT;
    cM = nM;
    cP = nP;
  }
  // Coarsest level: the inverse of its dense (n_top m)^2 system (the identity when the
  // only block is an identity T).  The info slots of levels this plan does not have are
  // zero.
  const int nt = p.n_top;
  const int Nt = nt * m;
  const long long nt2 = (long long)Nt * Nt;
  if (info) zero_int_async(info + info_base + 1, rb_info_count(nb) - info_base - 1, st);
  if (cT == nullptr) {
    if (info) zero_int_async(info + info_base, 1, st);
    return set_identity_batched(m, factor + p.top_off, m, m2, 1, st);
  }
  if (nt == 1)
    return inv_batched<E>(m, cT, m, m2, factor + p.top_off, m, m2, 1, LUw, ipiv,
                          info ? info + info_base : nullptr, 0.0, nullptr, st, inv_mode,
                          flags ? flags + info_base : nullptr);
  // nt > 1: red-black factor of the nt-node system (its blocks are the current level set,
  // stride m^2) down to one block -- its info entries fill this plan's remaining slots --
  // then its solve for the identity gives the inverse.
  const RbPlan sp = rb_plan(nt, m, 1);
  E* subF = Dw;
  E* subW = subF + sp.factor_doubles;
  E* Iw = subW + sp.work_doubles;
  SSLAB_TRY(rb_factor_t<E>(nt, m, cT, cM, cP, subF, subW, iwork, info ? info + info_base : nullptr,
                           st, 1, inv_mode, flags ? flags + info_base : nullptr));
  SSLAB_TRY(set_identity_batched(Nt, Iw, Nt, nt2, 1, st));
  return rb_solve_sub<E>(nt, m, Nt, subF, Iw, Nt, factor + p.top_off, Nt, st);
}

// The whole factorization with every inverse's backward-error check deferred: one
// synchronisation reads all the flags at the end, and if any inverse missed its
// tolerance the factorization is redone with pivoted-LU inverses throughout (every level
// depends on the ones before it).
template <typename E>
static int rb_factor_checked(int nb, int m, const E* T, const E* SiM, const E* SiP, E* factor,
                             E* work, int* iwork, int* info, hipStream_t st) {
  SSLAB_REQUIRE(nb >= 1 && m >= 1, "rb_factor: nb and m must be positive");
  SSLAB_REQUIRE(iwork, "rb_factor: null pointer");
  const RbPlan p = rb_plan(nb, m);
  const int nflag = rb_info_count(nb);
  int* flags = iwork + p.flag_off;
  zero_int_async(flags, nflag, st);
  SSLAB_TRY(rb_factor_t<E>(nb, m, T, SiM, SiP, factor, work, iwork, info, st, 0, kInvDeferred,
                           flags));
  std::vector<int> hf(nflag);
  SSLAB_HIP_CHECK(hipMemcpyAsync(hf.data(), flags, sizeof(int) * nflag, hipMemcpyDeviceToHost, st));
  SSLAB_HIP_CHECK(hipStreamSynchronize(st));
  for (int i = 0; i < nflag; ++i)
    if (hf[i] != 0)
      return rb_factor_t<E>(nb, m, T, SiM, SiP, factor, work, iwork, info, st, 0, kInvLU, nullptr);
  return kOk;
}

int rb_factor(int nb, int m, const double* T, const double* SiM, const double* SiP,
              double* factor, double* work, int* iwork, int* info, hipStream_t st) {
  return rb_factor_checked<double>(nb, m, T, SiM, SiP, factor, work, iwork, info, st);
}

int rb_factor_f32(int nb, int m, const float* T, const float* SiM, const float* SiP,
                  float* factor, float* work, int* iwork, int* info, hipStream_t st) {
  return rb_factor_checked<float>(nb, m, T, SiM, SiP, factor, work, iwork, info, st);
}

int rb_info_count(int nb) {
  int n = nb, c = 0;
  while (n > 1) {
    c += n / 2;
    n = (n + 1) / 2;
  }
  return c + 1;
}

// The solve sweeps, for factors (and right-hand sides) stored in FP64 or FP32.
// One system of a (possibly multi-system) red-black solve.
template <typename T>
struct RbSolveSys {
  int nb, m, r;
  const T* factor;
  T* V;
  int ldv;
  T* X;
  int ldx;
  RbPlan p;
  bool ident0;  // the first level's diagonal blocks are identities (T^{-1} = I there)
  bool rb0;     // entry state of oms_rhs_rb0 (requires ident0): V's even slots hold the
                // level-0 reduced v_i - SiM_i v_{i-1} - SiP_i v_{i+1}, X's odd slots y_j = v_j,
                // X's even slots zero; level 0's forward sweep and X's zeroing are skipped
  int phase;    // 0: whole solve; 1: forward sweep and coarsest level only, leaving in X
                // the backward sweep's entry state (y_j at every level's odd nodes, x at
                // the coarsest level's nodes); 2: backward sweep only, from that state
  int top_max;  // the plan's coarsest-level limit (0: the default, rb_top_max(m))
  bool chain;   // run the backward sweep as one dependency-ordered launch (FP64, one
                // system, at most kChainBlocks blocks); otherwise one launch per level
};

// Tile policies of the backward sweep's steps (step u runs level L-1-u), per calling
// thread; steps past the table use the thread's GEMM policy.  Empty by default.
constexpr int kMaxBwdPolicies = 16;
static thread_local GemmPolicy g_bwd_policy[kMaxBwdPolicies];
static thread_local int g_bwd_npolicy = 0;

int rb_set_backward_policies(const GemmPolicy* p, int n) {
  SSLAB_REQUIRE(n >= 0 && n <= kMaxBwdPolicies, "rb_set_backward_policies: 0..%d steps",
                kMaxBwdPolicies);
  SSLAB_REQUIRE(n == 0 || p != nullptr, "rb_set_backward_policies: null table");
  const GemmPolicy saved = gemm_get_policy();
  for (int u = 0; u < n; ++u) {  // validate every entry through the policy setter
    const int rc = gemm_set_policy(p[u]);
    if (rc != kOk) {
      gemm_set_policy(saved);
      return rc;
    }
  }
  gemm_set_policy(saved);
  for (int u = 0; u < n; ++u) g_bwd_policy[u] = p[u];
  g_bwd_npolicy = n;
  return kOk;
}

int rb_get_backward_policy_count() { return g_bwd_npolicy; }

// Append a descriptor, launching the pending ones first when the launch is full.  The
// descriptors of one level write disjoint outputs from inputs no descriptor of that
// level writes, so a level may be split over several launches.
template <typename T>
static int rb_push(GemmLaunchT<T>& G, const GemmDescT<T>& d, hipStream_t st) {
  if (G.nd == kMaxDesc) {
    SSLAB_TRY(gemm_launch(G, st));
    G.nd = 0;
  }
  G.d[G.nd++] = d;
  return kOk;
}

template <typename T>
static int rb_flush(GemmLaunchT<T>& G, hipStream_t st) {
  if (G.nd > 0) {
    SSLAB_TRY(gemm_launch(G, st));
    G.nd = 0;
  }
  return kOk;
}

// One system's products of a step: added to the step's shared launch G with the other
// systems' products when they have at most 32 columns (few right-hand sides: small
// launches that grouping fills out), else launched on their own.  With more columns a
// system's products already fill the chip or split K well by themselves, and a pooled
// launch picks its tiles and K split from the pooled counts, which splits the largest
// system's long K less and makes it the tail (n = 512 / 1024 / 2048 systems with 128
// right-hand sides: 13.4 ms pooled against 10.2 ms one system after another).
template <typename T>
static int rb_route(GemmLaunchT<T>& G, GemmLaunchT<T>& Gs, hipStream_t st) {
  if (Gs.nd == 0) return kOk;
  int nmax = 0;
  for (int i = 0; i < Gs.nd; ++i) nmax = Gs.d[i].N > nmax ? Gs.d[i].N : nmax;
  if (nmax > 32) return rb_flush(Gs, st);
  for (int i = 0; i < Gs.nd; ++i) SSLAB_TRY(rb_push(G, Gs.d[i], st));
  Gs.nd = 0;
  return kOk;
}

// Forward sweep of level l, first launch: y_j = Tinv_j v_j for odd j, accumulated into the
// zero-initialised X (with Tinv = I at an identity level 0, y_j = v_j is read from V by the
// even update and the backward sweep instead).
template <typename T>
static int rb_fwd_odd(GemmLaunchT<T>& G, const RbSolveSys<T>& S, int l, hipStream_t st) {
  if (l == 0 && S.ident0) return kOk;
  const int m = S.m, r = S.r, ldv = S.ldv, ldx = S.ldx;
  const long long m2 = (long long)m * m;
  const long long bv = (long long)m * ldv, bx = (long long)m * ldx;
  const RbLevel& v = S.p.lv[l];
  const T* TG = S.factor + v.tg_off;
  const long long sv = (2LL << l) * bv, sx = (2LL << l) * bx;  // same-parity node stride
  const long long ov = (1LL << l) * bv, ox = (1LL << l) * bx;  // first odd node
  GemmDescT<T> d = gemm_desc<T>(m, r, v.nO, S.X + ox, ldx, sx, S.X + ox, ldx, sx, 1.0);
  gemm_add_term<T>(d, gemm_term<T>(TG, 3 * m, 3 * m2, 0, S.V + ov, ldv, sv, 0, m, 1.0));
  return rb_push(G, d, st);
}

// Forward sweep of level l, second launch: v_i -= SiM_i y_{i-1} + SiP_i y_{i+1} (even i).
template <typename T>
static int rb_fwd_even(GemmLaunchT<T>& G, const RbSolveSys<T>& S, int l, hipStream_t st) {
  if (l == 0 && S.rb0) return kOk;  // the caller applied it with the rhs
  const int m = S.m, r = S.r, ldv = S.ldv;
  T* V = S.V;
  const long long m2 = (long long)m * m;
  const long long bv = (long long)m * ldv;
  const RbLevel& v = S.p.lv[l];
  const int nE = v.nE, nO = v.nO;
  const T* WW = S.factor + v.ww_off;
  const long long sv = (2LL << l) * bv, ov = (1LL << l) * bv;
  const bool from_v = l == 0 && S.ident0;
  const T* Y = from_v ? (const T*)V : (const T*)S.X;
  const int ldy = from_v ? ldv : S.ldx;
  const long long by = (long long)m * ldy;
  const long long sy = (2LL << l) * by, oy = (1LL << l) * by;
  const int both = (nE < nO ? nE : nO) - 1;
  if (both > 0) {  // even b in [1, min(nE,nO))
    GemmDescT<T> d = gemm_desc<T>(m, r, both, V + sv, ldv, sv, V + sv, ldv, sv, 1.0);
    gemm_add_term<T>(d, gemm_term<T>(WW + 2 * m2, 2 * m, 2 * m2, 0, Y + oy, ldy, sy, 0, m, -1.0));
    gemm_add_term<T>(d, gemm_term<T>(WW + 2 * m2 + m, 2 * m, 2 * m2, 0, Y + sy + oy, ldy, sy, 0, m,
                                     -1.0));
    SSLAB_TRY(rb_push(G, d, st));
  }
  {  // b = 0
    GemmDescT<T> d = gemm_desc<T>(m, r, 1, V, ldv, sv, V, ldv, sv, 1.0);
    gemm_add_term<T>(d, gemm_term<T>(WW + m, 2 * m, 2 * m2, 0, Y + oy, ldy, sy, 0, m, -1.0));
    SSLAB_TRY(rb_push(G, d, st));
  }
  if (nE > nO && nE - 1 >= 1) {
    const long long b = nE - 1;
    GemmDescT<T> d = gemm_desc<T>(m, r, 1, V + b * sv, ldv, sv, V + b * sv, ldv, sv, 1.0);
    gemm_add_term<T>(d, gemm_term<T>(WW + b * 2 * m2, 2 * m, 2 * m2, 0, Y + b * sy - oy, ldy, sy,
                                     0, m, -1.0));
    SSLAB_TRY(rb_push(G, d, st));
  }
  return kOk;
}

// Coarsest level: x_k = sum_j Inv_kj v_j over its n_top nodes (slots k << L), as
// descriptors batched over k with up to kMaxTerms summed j each.  X starts at zero there;
// with more than kMaxTerms nodes several descriptors add into every x_k, atomically.
template <typename T>
static int rb_top(GemmLaunchT<T>& G, const RbSolveSys<T>& S, hipStream_t st) {
  const int m = S.m, nt = S.p.n_top, L = S.p.L;
  const int Nt = nt * m;
  const long long sv = ((long long)m << L) * S.ldv, sx = ((long long)m << L) * S.ldx;
  const T* Inv = S.factor + S.p.top_off;
  for (int j0 = 0; j0 < nt; j0 += kMaxTerms) {
    GemmDescT<T> d = gemm_desc<T>(m, S.r, nt, S.X, S.ldx, sx, S.X, S.ldx, sx, 1.0);
    d.atomic = nt > kMaxTerms;
    for (int j = j0; j < nt && j < j0 + kMaxTerms; ++j)
      gemm_add_term<T>(d, gemm_term<T>(Inv + (long long)j * m, Nt, (long long)m * Nt, 0,
                                       S.V + j * sv, S.ldv, 0, 0, m, 1.0));
    SSLAB_TRY(rb_push(G, d, st));
  }
  return kOk;
}

// A backward-sweep descriptor with the row blocks of X it writes and reads (block i =
// rows [i m, (i + 1) m)), in the encoding of GemmChainT (index + 1, 0: none).
template <typename T>
struct BwdDesc {
  GemmDescT<T> d;
  int sig1, sig_stride;
  int dep1[kMaxTerms], dep_stride[kMaxTerms];
};

// Backward sweep of level l:  x_j = y_j - (Tinv SiM)_j x_{j-1} - (Tinv SiP)_j x_{j+1}.
// Odd node c of level l is block 2^l + c 2^(l+1); its neighbours are blocks c 2^(l+1)
// and (c + 1) 2^(l+1).  Returns the number of descriptors written to out (at most 2).
template <typename T>
static int rb_bwd_descs(const RbSolveSys<T>& S, int l, BwdDesc<T>* out) {
  const int m = S.m, r = S.r, ldx = S.ldx;
  T* X = S.X;
  const long long m2 = (long long)m * m;
  const long long bx = (long long)m * ldx;
  const RbLevel& v = S.p.lv[l];
  const int n = v.n, nO = v.nO;
  const T* TG = S.factor + v.tg_off;
  const long long sx = (2LL << l) * bx;
  const long long ox = (1LL << l) * bx;
  const int nboth = (n % 2 == 0) ? nO - 1 : nO;
  const bool from_v = (l == 0 && S.ident0 && !S.rb0 && S.phase == 0);
  const T* Y = from_v ? (const T*)S.V : (const T*)X;
  const int ldy = from_v ? S.ldv : ldx;
  const long long by = (long long)m * ldy;
  const long long sy = (2LL << l) * by, oy = (1LL << l) * by;
  const int half = 1 << l, step = 2 << l;
  int k = 0;
  if (nboth > 0) {
    BwdDesc<T>& e = out[k++];
    e = BwdDesc<T>{};
    e.d = gemm_desc<T>(m, r, nboth, X + ox, ldx, sx, Y + oy, ldy, sy, 1.0);
    gemm_add_term<T>(e.d, gemm_term<T>(TG + m, 3 * m, 3 * m2, 0, X, ldx, sx, 0, m, -1.0));
    gemm_add_term<T>(e.d, gemm_term<T>(TG + 2 * m, 3 * m, 3 * m2, 0, X + sx, ldx, sx, 0, m, -1.0));
    e.sig1 = half + 1;
    e.sig_stride = step;
    e.dep1[0] = 1;
    e.dep_stride[0] = step;
    e.dep1[1] = step + 1;
    e.dep_stride[1] = step;
  }
  if (nboth < nO) {  // last odd node: left neighbour only
    const long long c = nO - 1;
    BwdDesc<T>& e = out[k++];
    e = BwdDesc<T>{};
    e.d = gemm_desc<T>(m, r, 1, X + ox + c * sx, ldx, sx, Y + oy + c * sy, ldy, sy, 1.0);
    gemm_add_term<T>(e.d, gemm_term<T>(TG + c * 3 * m2 + m, 3 * m, 3 * m2, 0, X + c * sx, ldx,
                                       sx, 0, m, -1.0));
    e.sig1 = half + (int)c * step + 1;
    e.dep1[0] = (int)c * step + 1;
  }
  return k;
}

template <typename T>
static int rb_bwd_level(GemmLaunchT<T>& G, const RbSolveSys<T>& S, int l, hipStream_t st) {
  BwdDesc<T> e[2];
  const int k = rb_bwd_descs(S, l, e);
  for (int i = 0; i < k; ++i) SSLAB_TRY(rb_push(G, e[i].d, st));
  return kOk;
}

// Device counters of the dependency-ordered backward sweep, one set per GPU (allocated
// and cleared on first use; the kernel leaves them at zero).  Chained sweeps on one GPU
// must not overlap in time.
constexpr int kChainDevices = 64;
static std::mutex g_chain_mu;
static int* g_chain_mem[kChainDevices] = {};

static int chain_counters(int** out) {
  int dev = 0;
  SSLAB_HIP_CHECK(hipGetDevice(&dev));
  SSLAB_REQUIRE(dev >= 0 && dev < kChainDevices, "rb chain: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_chain_mu);
  if (g_chain_mem[dev] == nullptr) {
    int* p = nullptr;
    SSLAB_HIP_CHECK(hipMalloc(&p, sizeof(int) * (kChainBlocks + 4)));
    const hipError_t e = hipMemset(p, 0, sizeof(int) * (kChainBlocks + 4));
    if (e != hipSuccess) {
      (void)hipFree(p);
      SSLAB_HIP_CHECK(e);
    }
    g_chain_mem[dev] = p;
  }
  *out = g_chain_mem[dev];
  return kOk;
}

// Backward sweep of one system as a single dependency-ordered launch (k_gemm_chain):
// the levels' descriptors coarsest first, each tile waiting for the blocks it reads.
// *done stays false when the system does not fit a chain (more than kChainBlocks blocks
// or kChainDesc descriptors); the caller then runs the per-level launches.  The tile
// policy is the first backward-step policy when a table is set, else the thread's.
static int rb_bwd_chain(const RbSolveSys<double>& S, hipStream_t st, bool* done) {
  *done = false;
  if (S.nb > kChainBlocks || S.p.L < 1) return kOk;
  GemmChainT<double> C{};
  C.nd = 0;
  for (int u = 0; u < S.p.L; ++u) {
    BwdDesc<double> e[2];
    const int k = rb_bwd_descs(S, S.p.L - 1 - u, e);
    if (C.nd + k > kChainDesc) return kOk;
    for (int i = 0; i < k; ++i) {
      C.d[C.nd] = e[i].d;
      C.sig1[C.nd] = e[i].sig1;
      C.sig_stride[C.nd] = e[i].sig_stride;
      for (int t = 0; t < kMaxTerms; ++t) {
        C.dep1[C.nd][t] = e[i].dep1[t];
        C.dep_stride[C.nd][t] = e[i].dep_stride[t];
      }
      ++C.nd;
    }
  }
  int* mem = nullptr;
  SSLAB_TRY(chain_counters(&mem));
  C.cnt = mem;
  C.ctl = mem + kChainBlocks;
  C.spin_limit = 1 << 18;
  const GemmPolicy saved = gemm_get_policy();
  if (g_bwd_npolicy > 0) gemm_set_policy(g_bwd_policy[0]);
  const int rc = gemm_chain_launch<double>(C, st);
  if (g_bwd_npolicy > 0) gemm_set_policy(saved);
  SSLAB_TRY(rc);
  *done = true;
  return kOk;
}

int rb_chain_status(int* aborted) {
  SSLAB_REQUIRE(aborted != nullptr, "rb_chain_status: null output");
  int dev = 0;
  SSLAB_HIP_CHECK(hipGetDevice(&dev));
  SSLAB_REQUIRE(dev >= 0 && dev < kChainDevices, "rb chain: device %d out of range", dev);
  int* p = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_chain_mu);
    p = g_chain_mem[dev];
  }
  *aborted = 0;
  if (p == nullptr) return kOk;
  SSLAB_HIP_CHECK(hipDeviceSynchronize());
  SSLAB_HIP_CHECK(hipMemcpy(aborted, p + kChainBlocks + 2, sizeof(int), hipMemcpyDeviceToHost));
  if (*aborted) SSLAB_HIP_CHECK(hipMemset(p + kChainBlocks + 2, 0, sizeof(int)));
  return kOk;
}

// Solve several independent systems together: at every step the descriptors of all
// systems with at most 32 right-hand sides share one grouped GEMM launch (more when they
// exceed kMaxDesc); wider systems keep their own launches (rb_route).  Forward step
// t runs level t of every system that has it and the coarsest block of every system
// whose levels ended at t; backward step u runs level L_s - 1 - u of system s, so the
// systems finish together.  Every slot of X is written exactly once by a product that
// accumulates into X, which starts at zero (then split-K launches need no init pass).
template <typename T>
static int rb_solve_multi_t(int nsys, RbSolveSys<T>* sys, hipStream_t st) {
  int Lmax = 0;
  bool any = false;
  for (int s = 0; s < nsys; ++s) {
    RbSolveSys<T>& S = sys[s];
    SSLAB_REQUIRE(S.nb >= 1 && S.m >= 1 && S.r >= 0, "rb_solve: bad sizes (system %d)", s);
    S.p = rb_plan(S.nb, S.m, S.top_max);
    if (S.r == 0) continue;  // nothing to solve (empty buffers may be null)
    const bool uses_v = S.phase != 2;
    SSLAB_REQUIRE(S.ldx >= S.r && (!uses_v || S.ldv >= S.r), "rb_solve: leading dimension < nrhs");
    SSLAB_REQUIRE(S.factor && S.X && (!uses_v || (S.V && (void*)S.V != (void*)S.X)),
                  "rb_solve: null or aliased buffers");
    any = true;
    if (S.p.L > Lmax) Lmax = S.p.L;
    if (!S.rb0 && S.phase != 2) SSLAB_TRY(zero_2d<T>((long long)S.nb * S.m, S.r, S.X, S.ldx, st));
  }
  if (!any) return kOk;
  GemmLaunchT<T> G{};
  GemmLaunchT<T> Gs{};
  G.nd = 0;
  Gs.nd = 0;
  for (int t = 0; t <= Lmax; ++t) {
    // level t's odd products (and the coarsest blocks of the systems whose levels ended),
    // then, once they are complete, level t's even updates, which read their y_j
    for (int s = 0; s < nsys; ++s) {
      const RbSolveSys<T>& S = sys[s];
      if (S.r == 0 || S.phase == 2) continue;
      if (t < S.p.L) SSLAB_TRY(rb_fwd_odd(Gs, S, t, st));
      else if (t == S.p.L) SSLAB_TRY(rb_top(Gs, S, st));
      SSLAB_TRY(rb_route(G, Gs, st));
    }
    SSLAB_TRY(rb_flush(G, st));
    for (int s = 0; s < nsys; ++s) {
      const RbSolveSys<T>& S = sys[s];
      if (S.r == 0 || S.phase == 2 || t >= S.p.L) continue;
      SSLAB_TRY(rb_fwd_even(Gs, S, t, st));
      SSLAB_TRY(rb_route(G, Gs, st));
    }
    SSLAB_TRY(rb_flush(G, st));
  }
  for (int s = 0; s < nsys; ++s) {  // forward only: level 0's y_j = v_j (T^{-1} = I) into X
    const RbSolveSys<T>& S = sys[s];
    if (S.r == 0 || S.phase != 1 || !S.ident0 || S.rb0 || S.p.L == 0) continue;
    const long long bv = (long long)S.m * S.ldv, bx = (long long)S.m * S.ldx;
    SSLAB_TRY(copy_block_batched<T>(S.m, S.r, S.V + bv, S.ldv, 2 * bv, S.X + bx, S.ldx, 2 * bx,
                                    S.p.lv[0].nO, st));
  }
  if constexpr (std::is_same<T, double>::value) {
    if (nsys == 1 && sys[0].chain && sys[0].r > 0 && sys[0].phase != 1) {
      bool done = false;
      SSLAB_TRY(rb_bwd_chain(sys[0], st, &done));
      if (done) return kOk;
    }
  }
  const GemmPolicy saved = gemm_get_policy();
  for (int u = 0; u < Lmax; ++u) {
    if (u < g_bwd_npolicy) gemm_set_policy(g_bwd_policy[u]);
    int rc = kOk;
    for (int s = 0; s < nsys && rc == kOk; ++s) {
      const RbSolveSys<T>& S = sys[s];
      const int l = S.p.L - 1 - u;
      if (S.r == 0 || l < 0 || S.phase == 1) continue;
      rc = rb_bwd_level(Gs, S, l, st);
      if (rc == kOk) rc = rb_route(G, Gs, st);
    }
    if (rc == kOk) rc = rb_flush(G, st);
    if (u < g_bwd_npolicy) gemm_set_policy(saved);
    if (rc != kOk) return rc;
  }
  return kOk;
}

template <typename T>
static int rb_solve_sub(int nb, int m, int r, const T* factor, T* V, int ldv, T* X, int ldx,
                        hipStream_t st, int top_max) {
  RbSolveSys<T> S{nb, m, r, factor, V, ldv, X, ldx, RbPlan{}, false, false, 0, top_max};
  return rb_solve_multi_t<T>(1, &S, st);
}

template <typename T>
static int rb_solve_t(int nb, int m, int r, const T* factor, T* V, int ldv, T* X, int ldx,
                      hipStream_t st, bool ident0 = false, bool rb0 = false, int phase = 0,
                      bool chain = false) {
  RbSolveSys<T> S{nb, m, r, factor, V, ldv, X, ldx, RbPlan{}, ident0, rb0, phase, 0, chain};
  return rb_solve_multi_t<T>(1, &S, st);
}

int rb_solve_ex(int nb, int m, int r, const double* factor, double* V, int ldv, double* X,
                int ldx, int flags, hipStream_t st) {
  SSLAB_REQUIRE((flags & ~31) == 0, "rb_solve: unknown flags 0x%x", flags);
  SSLAB_REQUIRE(!(flags & 2) || ((flags & 1) && nb >= 2),
                "rb_solve: the level-0-reduced entry state needs identity diagonal blocks and "
                "nb >= 2");
  SSLAB_REQUIRE((flags & 12) != 12, "rb_solve: forward-only and backward-only exclude each other");
  SSLAB_REQUIRE(!((flags & 2) && (flags & 8)),
                "rb_solve: backward-only already starts from the forward sweep's result");
  const int phase = (flags & 4) ? 1 : ((flags & 8) ? 2 : 0);
  return rb_solve_t<double>(nb, m, r, factor, V, ldv, X, ldx, st, (flags & 1) != 0,
                            (flags & 2) != 0, phase, (flags & 16) != 0);
}

// Forward sweep and coarsest level of the identity-diagonal red-black solve for a V whose
// block row i is zero outside the columns [lo[i], hi[i]) (host arrays): every product reads
// and writes only the columns where its operands can be nonzero -- an odd node's y_j keeps
// the support of v_j, an even node's update takes the union of its own and its
// neighbours' supports, and the coarsest level adds each Inv_kj v_j into x_k over v_j's
// support.  X (zeroed here) receives the backward sweep's entry state, exactly as
// rb_solve_ex with the identity-diagonal and forward-only flags leaves it; V is
// overwritten.  The boundary-map fold runs this on the rhs maps of the identity, whose
// block rows read a few boundary windows each.
int rb_forward_windowed(int nb, int m, int ncols, const double* factor, double* V, int ldv,
                        double* X, int ldx, const int* lo, const int* hi, hipStream_t st) {
  SSLAB_REQUIRE(nb >= 2 && m >= 1 && ncols >= 1, "rb_forward_windowed: bad sizes");
  SSLAB_REQUIRE(factor && V && X && (void*)V != (void*)X && lo && hi && ldv >= ncols &&
                    ldx >= ncols,
                "rb_forward_windowed: bad pointers / strides");
  std::vector<int> slo(nb), shi(nb);
  for (int i = 0; i < nb; ++i) {
    SSLAB_REQUIRE(lo[i] >= 0 && lo[i] < hi[i] && hi[i] <= ncols,
                  "rb_forward_windowed: block %d support [%d, %d) outside [0, %d)", i, lo[i],
                  hi[i], ncols);
    slo[i] = lo[i];
    shi[i] = hi[i];
  }
  const RbPlan p = rb_plan(nb, m);
  const long long m2 = (long long)m * m;
  const long long bv = (long long)m * ldv, bx = (long long)m * ldx;
  SSLAB_TRY(zero_2d<double>((long long)nb * m, ncols, X, ldx, st));
  GemmLaunch G{};
  G.nd = 0;
  for (int l = 0; l < p.L; ++l) {
    const RbLevel& v = p.lv[l];
    const double* TG = factor + v.tg_off;
    const double* WW = factor + v.ww_off;
    if (l > 0) {  // y_j = Tinv_j v_j (level 0: T = I, y_j = v_j is read from V)
      for (int c = 0; c < v.nO; ++c) {
        const long long j = (long long)(2 * c + 1) << l;
        const int w = shi[j] - slo[j];
        GemmDesc d = gemm_desc(m, w, 1, X + j * bx + slo[j], ldx, 0, X + j * bx + slo[j], ldx, 0,
                               1.0);
        gemm_add_term(d, gemm_term(TG + (long long)c * 3 * m2, 3 * m, 0, 0, V + j * bv + slo[j],
                                   ldv, 0, 0, m, 1.0));
        SSLAB_TRY(rb_push(G, d, st));
      }
      SSLAB_TRY(rb_flush(G, st));
    }
    const double* Y = l == 0 ? (const double*)V : (const double*)X;
    const long long by = l == 0 ? bv : bx;
    const int ldy = l == 0 ? ldv : ldx;
    for (int b = 0; b < v.nE; ++b) {  // v_i -= SiM_i y_{i-1} + SiP_i y_{i+1}
      const long long i = (long long)(2 * b) << l;
      const bool left = b >= 1, right = b < v.nO;
      const long long jl = i - (1LL << l), jr = i + (1LL << l);
      int nlo = slo[i], nhi = shi[i];
      if (left) nlo = std::min(nlo, slo[jl]), nhi = std::max(nhi, shi[jl]);
      if (right) nlo = std::min(nlo, slo[jr]), nhi = std::max(nhi, shi[jr]);
      GemmDesc d = gemm_desc(m, nhi - nlo, 1, V + i * bv + nlo, ldv, 0, V + i * bv + nlo, ldv, 0,
                             1.0);
      if (left)
        gemm_add_term(d, gemm_term(WW + (long long)b * 2 * m2, 2 * m, 0, 0, Y + jl * by + nlo, ldy,
                                   0, 0, m, -1.0));
      if (right)
        gemm_add_term(d, gemm_term(WW + (long long)b * 2 * m2 + m, 2 * m, 0, 0, Y + jr * by + nlo,
                                   ldy, 0, 0, m, -1.0));
      SSLAB_TRY(rb_push(G, d, st));
      slo[i] = nlo;
      shi[i] = nhi;
    }
    SSLAB_TRY(rb_flush(G, st));
  }
  // level 0's y_j = v_j into X (zero outside the supports, so whole rows are copied)
  SSLAB_TRY(copy_block_batched<double>(m, ncols, V + bv, ldv, 2 * bv, X + bx, ldx, 2 * bx,
                                       p.lv[0].nO, st));
  // coarsest level: x_k += Inv_kj v_j over v_j's support, atomically when several j
  const int nt = p.n_top, L = p.L;
  const long long Nt = (long long)nt * m;
  const double* Inv = factor + p.top_off;
  for (int k = 0; k < nt; ++k) {
    const long long sk = (long long)k << L;
    for (int j = 0; j < nt; ++j) {
      const long long sj = (long long)j << L;
      const int w = shi[sj] - slo[sj];
      GemmDesc d = gemm_desc(m, w, 1, X + sk * bx + slo[sj], ldx, 0, X + sk * bx + slo[sj], ldx, 0,
                             1.0);
      d.atomic = nt > 1;
      gemm_add_term(d, gemm_term(Inv + (long long)k * m * Nt + (long long)j * m, (int)Nt, 0, 0,
                                 V + sj * bv + slo[sj], ldv, 0, 0, m, 1.0));
      SSLAB_TRY(rb_push(G, d, st));
    }
  }
  return rb_flush(G, st);
}

int rb_solve_multi(int nsys, const int* nb, const int* m, const int* r, const double* const* factor,
                   double* const* V, const int* ldv, double* const* X, const int* ldx,
                   hipStream_t st) {
  SSLAB_REQUIRE(nsys >= 0 && nsys <= 64, "rb_solve_multi: 0 <= nsys <= 64");
  if (nsys == 0) return kOk;
  SSLAB_REQUIRE(nb && m && r && factor && V && ldv && X && ldx, "rb_solve_multi: null array");
  std::vector<RbSolveSys<double>> sys(nsys);
  for (int s = 0; s < nsys; ++s)
    sys[s] = RbSolveSys<double>{nb[s], m[s], r[s], factor[s], V[s], ldv[s], X[s], ldx[s], RbPlan{},
                                false, false, 0, 0};
  return rb_solve_multi_t<double>(nsys, sys.data(), st);
}

int rb_solve(int nb, int m, int r, const double* factor, double* V, int ldv, double* X, int ldx,
             hipStream_t st) {
  return rb_solve_t<double>(nb, m, r, factor, V, ldv, X, ldx, st);
}

int rb_solve_f32(int nb, int m, int r, const float* factor, float* V, int ldv, float* X, int ldx,
                 hipStream_t st) {
  return rb_solve_t<float>(nb, m, r, factor, V, ldv, X, ldx, st);
}

// r = b - S x for the block system (T = NULL means identity diagonal blocks), FP64.
int rb_residual(int nb, int m, int r, const double* T, const double* SiM, const double* SiP,
                const double* B, int ldb, const double* X, int ldx, double* R, int ldr,
                hipStream_t st) {
  SSLAB_REQUIRE(nb >= 1 && m >= 1 && r >= 0 && SiM && SiP && B && X && R,
                "rb_residual: bad arguments");
  if (r == 0) return kOk;
  const long long m2 = (long long)m * m;
  const long long bb = (long long)m * ldb, bx = (long long)m * ldx, br = (long long)m * ldr;
  GemmLaunch G{};
  int k = 0;
  const double* Cin = B;
  int ldcin = ldb;
  long long scin = bb;
  if (T == nullptr) {  // R = B - X first, then the couplings accumulate into R
    SSLAB_REQUIRE(ldb == r && ldx == r && ldr == r, "rb_residual: contiguous buffers required");
    SSLAB_TRY(sub_f64((long long)nb * m * r, B, X, R, st));
    Cin = R;
    ldcin = ldr;
    scin = br;
  }
  auto add_diag = [&](GemmDesc& d, long long i0) {
    if (T) gemm_add_term(d, gemm_term(T + i0 * m2, m, m2, 0, X + i0 * bx, ldx, bx, 0, m, -1.0));
  };
  if (nb >= 3) {  // interior blocks: both neighbours
    GemmDesc d = gemm_desc(m, r, nb - 2, R + br, ldr, br, Cin + scin, ldcin, scin, 1.0);
    gemm_add_term(d, gemm_term(SiM + m2, m, m2, 0, X, ldx, bx, 0, m, -1.0));
    gemm_add_term(d, gemm_term(SiP + m2, m, m2, 0, X + 2 * bx, ldx, bx, 0, m, -1.0));
    add_diag(d, 1);
    G.d[k++] = d;
  }
  {
    GemmDesc d = gemm_desc(m, r, 1, R, ldr, br, Cin, ldcin, scin, 1.0);
    if (nb >= 2) gemm_add_term(d, gemm_term(SiP, m, m2, 0, X + bx, ldx, bx, 0, m, -1.0));
    add_diag(d, 0);
    if (d.nterms) G.d[k++] = d;
  }
  if (nb >= 2) {
    const long long i = nb - 1;
    GemmDesc d = gemm_desc(m, r, 1, R + i * br, ldr, br, Cin + i * scin, ldcin, scin, 1.0);
    gemm_add_term(d, gemm_term(SiM + i * m2, m, m2, 0, X + (i - 1) * bx, ldx, bx, 0, m, -1.0));
    add_diag(d, i);
    G.d[k++] = d;
  }
  G.nd = k;
  if (k) SSLAB_TRY(gemm_launch(G, st));
  return kOk;
}

// Mixed-precision solve: the red-black factor stored in FP32 is the preconditioner of
// an iterative refinement whose residual r = b - S x is FP64:
//   x = M b;  repeat `iters` times:  r = b - S x (FP64),  x += M r.
// B, X: N x r (contiguous, N = nb*m); R64: N x r doubles; V32, X32: N x r floats each.
int rb_refine_mixed(int nb, int m, int r, const float* factor32, const double* T,
                    const double* SiM, const double* SiP, const double* B, double* X,
                    double* R64, float* V32, float* X32, int iters, hipStream_t st) {
  SSLAB_REQUIRE(nb >= 1 && m >= 1 && r >= 0 && iters >= 0, "rb_refine_mixed: bad sizes");
  if (r == 0 || iters == 0) return kOk;
  SSLAB_REQUIRE(factor32 && B && X && V32 && X32 && R64 && SiM && SiP,
                "rb_refine_mixed: null pointer");
  const long long n = (long long)nb * m * r;
  for (int it = 0; it < iters; ++it) {
    SSLAB_TRY(rb_residual(nb, m, r, T, SiM, SiP, B, r, X, r, R64, r, st));
    SSLAB_TRY(cast_f64_to_f32(n, R64, V32, st));
    SSLAB_TRY(rb_solve_t<float>(nb, m, r, factor32, V32, r, X32, r, st));
    SSLAB_TRY(acc_f32_into_f64(n, X32, X, st));
  }
  return kOk;
}

int rb_solve_mixed(int nb, int m, int r, const float* factor32, const double* T,
                   const double* SiM, const double* SiP, const double* B, double* X, double* R64,
                   float* V32, float* X32, int iters, hipStream_t st) {
  SSLAB_REQUIRE(nb >= 1 && m >= 1 && r >= 0 && iters >= 0, "rb_solve_mixed: bad sizes");
  SSLAB_REQUIRE(factor32 && B && X && V32 && X32 && (iters == 0 || (R64 && SiM && SiP)),
                "rb_solve_mixed: null pointer");
  if (r == 0) return kOk;
  const long long n = (long long)nb * m * r;
  SSLAB_TRY(cast_f64_to_f32(n, B, V32, st));
  SSLAB_TRY(rb_solve_t<float>(nb, m, r, factor32, V32, r, X32, r, st));
  SSLAB_TRY(cast_f32_to_f64(n, X32, X, st));
  return rb_refine_mixed(nb, m, r, factor32, T, SiM, SiP, B, X, R64, V32, X32, iters, st);
}

int rb_solve_f32_ex(int nb, int m, int r, const float* factor, float* V, int ldv, float* X,
                    int ldx, int flags, hipStream_t st) {
  SSLAB_REQUIRE((flags & ~15) == 0, "rb_solve_f32: unknown flags 0x%x", flags);
  SSLAB_REQUIRE(!(flags & 2) || ((flags & 1) && nb >= 2),
                "rb_solve_f32: the level-0-reduced entry state needs identity diagonal blocks "
                "and nb >= 2");
  SSLAB_REQUIRE((flags & 12) != 12,
                "rb_solve_f32: forward-only and backward-only exclude each other");
  SSLAB_REQUIRE(!((flags & 2) && (flags & 8)),
                "rb_solve_f32: backward-only already starts from the forward sweep's result");
  const int phase = (flags & 4) ? 1 : ((flags & 8) ? 2 : 0);
  return rb_solve_t<float>(nb, m, r, factor, V, ldv, X, ldx, st, (flags & 1) != 0,
                           (flags & 2) != 0, phase);
}

// ---------------------------------------------------------------------------
// rb_solve_mixed without a caller workspace: the FP64 residual and the two FP32 panels
// live in a file-scope device array (one per GPU, held until the process ends).  Host
// code never passes the array's name to a kernel: its device address is looked up once
// per device with hipGetSymbolAddress and cached here.

namespace {

constexpr long long kMixedScratchBytes = 64LL << 20;
__device__ __align__(256) unsigned char scratch[kMixedScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

long long round256(long long b) { return (b + 255) / 256 * 256; }

// Device address of this GPU's scratch array (looked up on the first call per device).
int mixed_scratch(void** out) {
  int dev = 0;
  SSLAB_HIP_CHECK(hipGetDevice(&dev));
  SSLAB_REQUIRE(dev >= 0 && dev < kMaxDevices, "rb_solve_mixed_noalloc: device %d out of range",
                dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    SSLAB_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SSLAB_REQUIRE(p != nullptr, "rb_solve_mixed_noalloc: no device address for the scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

}  // namespace

int rb_solve_mixed_noalloc(int nb, int m, int r, const float* factor32, const double* T,
                           const double* SiM, const double* SiP, const double* B, double* X,
                           int iters, hipStream_t st) {
  SSLAB_REQUIRE(nb >= 1 && m >= 1 && r >= 0 && iters >= 0,
                "rb_solve_mixed_noalloc: bad sizes");
  const long long n = (long long)nb * m * r;
  const long long need = round256(8 * n) + 2 * round256(4 * n);
  SSLAB_REQUIRE(need <= kMixedScratchBytes,
                "rb_solve_mixed_noalloc: nb*m*nrhs = %lld needs %lld bytes of scratch, more "
                "than the %lld this library holds per GPU (call sslab_rb_solve_mixed with a "
                "workspace instead)",
                n, need, kMixedScratchBytes);
  if (r == 0) return kOk;
  void* base = nullptr;
  SSLAB_TRY(mixed_scratch(&base));
  unsigned char* p = static_cast<unsigned char*>(base);
  double* R64 = reinterpret_cast<double*>(p);
  float* V32 = reinterpret_cast<float*>(p + round256(8 * n));
  float* X32 = reinterpret_cast<float*>(p + round256(8 * n) + round256(4 * n));
  // every panel is written before it is read (cast, solve, residual), so nothing of an
  // earlier call's contents is used
  return rb_solve_mixed(nb, m, r, factor32, T, SiM, SiP, B, X, R64, V32, X32, iters, st);
}

// ---------------------------------------------------------------------------
// Thomas (sequential) variant.
// factor layout: A (nb-1 blocks) | Binv (nb) | LU (nb) | C (nb-1);  ipiv: nb*m ints.

int thomas_sizes(int nb, int m, long long* factor_doubles, long long* work_doubles,
                 long long* work_ints) {
  SSLAB_REQUIRE(nb >= 1 && m >= 1, "thomas_sizes: nb and m must be positive");
  const long long m2 = (long long)m * m;
  if (factor_doubles) *factor_doubles = (4LL * nb - 2) * m2;
  if (work_doubles) *work_doubles = m2;
  if (work_ints) *work_ints = (long long)nb * m + m;
  return kOk;
}

int thomas_factor(int nb, int m, const double* T, const double* SiM, const double* SiP,
                  double* factor, double* work, int* iwork, int* info, hipStream_t st) {
  SSLAB_REQUIRE(nb >= 1 && m >= 1, "thomas_factor: nb and m must be positive");
  SSLAB_REQUIRE(T && factor && work && iwork, "thomas_factor: null pointer");
  const long long m2 = (long long)m * m;
  double* A = factor;
  double* Binv = A + (long long)(nb - 1) * m2;
  double* LU = Binv + (long long)nb * m2;
  double* C = LU + (long long)nb * m2;
  int* ipiv = iwork;
  int* perm = iwork + (long long)nb * m;
  if (nb > 1) SSLAB_TRY(copy_block_batched(m, m, SiP, m, m2, C, m, m2, nb - 1, st));
  SSLAB_TRY(copy_block_batched(m, m, T, m, m2, LU, m, m2, 1, st));
  for (int i = 0; i < nb; ++i) {
    double* LUi = LU + (long long)i * m2;
    if (i > 0) {
      // B_i = T_i - A_i C_{i-1}, with A_i = SiM_i Binv_{i-1}.
      double* Ai = A + (long long)(i - 1) * m2;
      SSLAB_TRY(gemm_strided(0, 0, m, m, m, 1.0, SiM + (long long)i * m2, m, 0,
                             Binv + (long long)(i - 1) * m2, m, 0, 0.0, Ai, m, 0, 1, st));
      GemmLaunch G{};
      G.nd = 1;
      G.d[0] = gemm_desc(m, m, 1, LUi, m, 0, T + (long long)i * m2, m, 0, 1.0);
      gemm_add_term(G.d[0], gemm_term(Ai, m, 0, 0, C + (long long)(i - 1) * m2, m, 0, 0, m, -1.0));
      SSLAB_TRY(gemm_launch(G, st));
    }
    SSLAB_TRY(getrf_batched(m, LUi, m, m2, ipiv + (long long)i * m, m, info ? info + i : nullptr,
                            1, st));
    SSLAB_TRY(set_identity_batched(m, work, m, m2, 1, st));
    SSLAB_TRY(getrs_batched(m, m, LUi, m, m2, ipiv + (long long)i * m, m, work, m, m2,
                            Binv + (long long)i * m2, m, m2, perm, 1, st));
  }
  return kOk;
}

int thomas_solve(int nb, int m, int r, const double* factor, double* V, int ldv, double* X,
                 int ldx, hipStream_t st) {
  SSLAB_REQUIRE(nb >= 1 && m >= 1 && r >= 0, "thomas_solve: bad sizes");
  SSLAB_REQUIRE(ldv >= r && ldx >= r, "thomas_solve: leading dimension < nrhs");
  SSLAB_REQUIRE(factor && V && X && V != X, "thomas_solve: null or aliased buffers");
  if (r == 0) return kOk;
  const long long m2 = (long long)m * m;
  const double* A = factor;
  const double* Binv = A + (long long)(nb - 1) * m2;
  const double* C = Binv + 2LL * nb * m2;
  const long long bv = (long long)m * ldv, bx = (long long)m * ldx;
  // forward: d_i -= A_i d_{i-1}
  for (int i = 1; i < nb; ++i)
    SSLAB_TRY(gemm_strided(0, 0, m, r, m, -1.0, A + (long long)(i - 1) * m2, m, 0,
                           V + (long long)(i - 1) * bv, ldv, 0, 1.0, V + (long long)i * bv, ldv, 0,
                           1, st));
  // backward: x_{n-1} = Binv d_{n-1};  x_i = Binv_i (d_i - C_i x_{i+1})
  SSLAB_TRY(gemm_strided(0, 0, m, r, m, 1.0, Binv + (long long)(nb - 1) * m2, m, 0,
                         V + (long long)(nb - 1) * bv, ldv, 0, 0.0, X + (long long)(nb - 1) * bx,
                         ldx, 0, 1, st));
  for (int i = nb - 2; i >= 0; --i) {
    SSLAB_TRY(gemm_strided(0, 0, m, r, m, -1.0, C + (long long)i * m2, m, 0,
                           X + (long long)(i + 1) * bx, ldx, 0, 1.0, V + (long long)i * bv, ldv, 0,
                           1, st));
    SSLAB_TRY(gemm_strided(0, 0, m, r, m, 1.0, Binv + (long long)i * m2, m, 0,
                           V + (long long)i * bv, ldv, 0, 0.0, X + (long long)i * bx, ldx, 0, 1,
                           st));
  }
  return kOk;
}

}  // namespace sslab
