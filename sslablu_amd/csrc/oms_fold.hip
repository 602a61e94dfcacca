// Boundary-folded forward sweep of the oms S-system solve (FP64, gfx950).
//
// Everything the red-black solve computes before its backward sweep is linear in the
// global boundary data F (nF rows ordered [bottom | top | left | right]): the backward
// sweep's entry state E -- y_j = T_j^{-1} v_j at every level's odd nodes and x at the
// coarsest level's nodes, one m-row block per interface -- equals M F for a single
// (nb m) x nF matrix M.  The block of M of an odd node of level l reads only the bottom
// and top windows of the few double slabs its eliminations touched (about 2^(l+2)
// windows of nx*p rows), a few hundred of the nF columns, plus the left / right boundary
// data for the nodes at the two ends.
//
// M is built once per operator by running the rhs maps and the forward sweep on F = I
// (its columns).  Its nonzero column runs -- at most kMaxTerms per block after closing
// small gaps -- are gathered into compact arrays, one per group, where a group is a
// sequence of same-level blocks whose runs have equal lengths and shift by a constant
// number of columns from one block to the next.  A solve then is one zero kernel and one
// grouped GEMM launch, X_b = sum_t Mc_{b,t} F[run_t], ahead of the backward sweep: the
// forward sweep costs O(window) instead of O(m) columns per block and no longer runs
// level by level.
#include "oms_fold.hpp"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <vector>

#include "blocktri_f64.hpp"
#include "gemm.hpp"
#include "lu_f64.hpp"

namespace sslab {

struct FoldTerm {
  long long a_off;  // compact map of the group's first block (offset into maps)
  int lda;
  long long sA;     // next block's map
  int f_row;        // first row of F read by the group's first block
  int f_rstride;    // shift of that row from one block to the next
  int K;
};

struct FoldGroup {
  int batch;      // blocks in the group
  int x_row;      // first row of X written (block slot * m)
  int x_rstride;  // rows from one block to the next
  int nterms;
  FoldTerm t[kMaxTerms];
};

struct FoldPlan {
  int nb, m, nF;
  double* maps;
  long long maps_doubles;
  std::vector<FoldGroup> groups;
  int* zero_rows;    // device: first X row of every block of the split groups
  int nzero;         // number of such blocks
  int nosplit_k;     // groups whose summed K is below this are never split
  float* maps32;     // the maps rounded to FP32 (oms_fold_enable_f32), or null
};

// Groups whose summed K is below this are never split: their blocks are written once
// with plain stores, so only the blocks of the long groups need zeroed rows.  The
// environment variable SSLAB_FOLD_NOSPLIT_K overrides it (read when a plan is built).
constexpr int kFoldNoSplitK = 512;

int fold_nosplit_k() {
  static const int v = [] {
    const char* e = getenv("SSLAB_FOLD_NOSPLIT_K");
    const int x = e ? atoi(e) : 0;
    return x > 0 ? x : kFoldNoSplitK;
  }();
  return v;
}

namespace {

// Zero rows [row_b, row_b + m) x [0, r) of X for every listed block b (one launch).
template <typename T>
__global__ void k_zero_blocks(int nblk, const int* __restrict__ rows, int m, int r,
                              T* __restrict__ X, int ldx) {
  const long long n = (long long)nblk * m * r;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const long long q = e / r;
    const int c = (int)(e - q * r);
    const int b = (int)(q / m), i = (int)(q - (long long)b * m);
    X[(long long)(rows[b] + i) * ldx + c] = T(0);
  }
}

// flag[b * nF + c] = 1 when column c of block b (rows [b m, (b+1) m)) has a nonzero.
__global__ void k_col_nonzero(int nb, int m, int nF, const double* __restrict__ M, int ldm,
                              int* __restrict__ flag) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= (long long)nb * nF) return;
  const int b = (int)(e / nF);
  const int c = (int)(e - (long long)b * nF);
  const double* col = M + (long long)b * m * ldm + c;
  int nz = 0;
  for (int i = 0; i < m; ++i) nz |= (col[(long long)i * ldm] != 0.0) ? 1 : 0;
  flag[e] = nz;
}

struct Run {
  int s, e;
};

// Nonzero column runs of one block: aligned to even columns (16-byte operand loads),
// gaps of at most 8 columns closed, then the two closest runs merged until at most
// kMaxTerms remain (zero columns inside a run cost only multiply-adds of zeros).
std::vector<Run> block_runs(const int* f, int nF) {
  std::vector<Run> r;
  for (int c = 0; c < nF;) {
    if (!f[c]) {
      ++c;
      continue;
    }
    int e = c;
    while (e < nF && f[e]) ++e;
    r.push_back({c & ~1, std::min(nF, (e + 1) & ~1)});
    c = e;
  }
  std::vector<Run> o;
  for (const Run& x : r) {
    if (!o.empty() && x.s <= o.back().e + 8)
      o.back().e = std::max(o.back().e, x.e);
    else
      o.push_back(x);
  }
  while ((int)o.size() > kMaxTerms) {
    size_t best = 0;
    int gap = INT_MAX;
    for (size_t i = 0; i + 1 < o.size(); ++i) {
      if (o[i + 1].s - o[i].e < gap) {
        gap = o[i + 1].s - o[i].e;
        best = i;
      }
    }
    o[best].e = o[best + 1].e;
    o.erase(o.begin() + (long)best + 1);
  }
  return o;
}

bool same_lengths(const std::vector<Run>& a, const std::vector<Run>& b) {
  if (a.size() != b.size()) return false;
  for (size_t t = 0; t < a.size(); ++t)
    if (a[t].e - a[t].s != b[t].e - b[t].s) return false;
  return true;
}

}  // namespace

int oms_fold_create(int nb, int m, int nF, const double* maps, int ldm, FoldPlan** out,
                    hipStream_t st) {
  SSLAB_REQUIRE(out != nullptr, "oms_fold_create: null plan pointer");
  *out = nullptr;
  SSLAB_REQUIRE(nb >= 1 && m >= 1 && nF >= 1 && maps && ldm >= nF, "oms_fold_create: bad sizes");
  SSLAB_REQUIRE((long long)nb * m <= INT_MAX && (long long)nb * nF <= INT_MAX,
                "oms_fold_create: problem too large");
  int nodd[64];
  int n_top = 0;
  const int L = rb_shape(nb, m, nodd, 64, &n_top);
  SSLAB_REQUIRE(L >= 0 && L < 64, "oms_fold_create: bad level count");

  // 1. nonzero columns of every block, to the host
  const long long nfl = (long long)nb * nF;
  int* dflag = nullptr;
  SSLAB_HIP_CHECK(hipMalloc(&dflag, sizeof(int) * nfl));
  hipLaunchKernelGGL(k_col_nonzero, dim3((unsigned)((nfl + 255) / 256)), dim3(256), 0, st, nb, m,
                     nF, maps, ldm, dflag);
  std::vector<int> flag(nfl);
  hipError_t e1 = hipGetLastError();
  hipError_t e2 = e1 == hipSuccess
                      ? hipMemcpyAsync(flag.data(), dflag, sizeof(int) * nfl, hipMemcpyDeviceToHost, st)
                      : e1;
  hipError_t e3 = e2 == hipSuccess ? hipStreamSynchronize(st) : e2;
  (void)hipFree(dflag);
  SSLAB_HIP_CHECK(e3);

  std::vector<std::vector<Run>> runs(nb);
  for (int b = 0; b < nb; ++b) runs[b] = block_runs(flag.data() + (long long)b * nF, nF);

  // 2. blocks in level order (odd nodes of level 0, 1, ..., then the coarsest level),
  //    cut into groups of equal run lengths and constant run shifts
  std::vector<std::vector<int>> seqs;
  std::vector<int> seq_stride;  // slots between consecutive blocks of a sequence
  for (int l = 0; l < L; ++l) {
    std::vector<int> s;
    for (int c = 0; c < nodd[l]; ++c) s.push_back((2 * c + 1) << l);
    seqs.push_back(s);
    seq_stride.push_back(2 << l);
  }
  {
    std::vector<int> s;
    for (int c = 0; c < n_top; ++c) s.push_back(c << L);
    seqs.push_back(s);
    seq_stride.push_back(1 << L);
  }
  struct Proto {
    int b0, count, stride;
    std::vector<Run> r;
    std::vector<int> d;
  };
  std::vector<Proto> protos;
  for (size_t q = 0; q < seqs.size(); ++q) {
    const std::vector<int>& s = seqs[q];
    size_t i = 0;
    while (i < s.size()) {
      Proto p{s[i], 1, seq_stride[q], runs[s[i]], std::vector<int>(runs[s[i]].size(), 0)};
      size_t j = i + 1;
      while (j < s.size() && same_lengths(runs[s[j]], p.r)) {
        bool ok = true;
        for (size_t t = 0; t < p.r.size(); ++t) {
          const int d = runs[s[j]][t].s - runs[s[j - 1]][t].s;
          if (j == i + 1)
            p.d[t] = d;
          else if (d != p.d[t])
            ok = false;
        }
        if (!ok) break;
        ++p.count;
        ++j;
      }
      if (!p.r.empty()) protos.push_back(p);  // a block without runs stays zero
      i = j;
    }
  }

  // 3. compact maps: group g holds count blocks of m x Ktot (its runs side by side)
  FoldPlan* P = new FoldPlan{nb, m, nF, nullptr, 0, {}, nullptr, 0, kFoldNoSplitK, nullptr};
  long long total = 0;
  for (const Proto& p : protos) {
    int kt = 0;
    for (const Run& x : p.r) kt += x.e - x.s;
    total += (long long)p.count * m * kt;
  }
  P->maps_doubles = total;
  if (total > 0) {
    hipError_t ea = hipMalloc(&P->maps, sizeof(double) * total);
    if (ea != hipSuccess) {
      delete P;
      SSLAB_HIP_CHECK(ea);
    }
  }
  long long off = 0;
  for (const Proto& p : protos) {
    int kt = 0;
    for (const Run& x : p.r) kt += x.e - x.s;
    FoldGroup g{};
    g.batch = p.count;
    g.x_row = p.b0 * m;
    g.x_rstride = p.stride * m;
    g.nterms = (int)p.r.size();
    int coff = 0;
    for (int t = 0; t < g.nterms; ++t) {
      const int len = p.r[t].e - p.r[t].s;
      g.t[t] = FoldTerm{off + coff, kt, (long long)m * kt, p.r[t].s, p.d[t], len};
      const int rc = copy_block_batched<double>(
          m, len, maps + (long long)p.b0 * m * ldm + p.r[t].s, ldm,
          (long long)p.stride * m * ldm + p.d[t], P->maps + off + coff, kt, (long long)m * kt,
          p.count, st);
      if (rc != kOk) {
        oms_fold_destroy(P);
        return rc;
      }
      coff += len;
    }
    P->groups.push_back(g);
    off += (long long)p.count * m * kt;
  }
  // the blocks of split-eligible groups start from zero rows (see kFoldNoSplitK)
  P->nosplit_k = fold_nosplit_k();
  std::vector<int> zr;
  for (const FoldGroup& g : P->groups) {
    int kt = 0;
    for (int t = 0; t < g.nterms; ++t) kt += g.t[t].K;
    if (kt >= P->nosplit_k)
      for (int q = 0; q < g.batch; ++q) zr.push_back(g.x_row + q * g.x_rstride);
  }
  P->nzero = (int)zr.size();
  if (!zr.empty()) {
    hipError_t ez = hipMalloc(&P->zero_rows, sizeof(int) * zr.size());
    if (ez == hipSuccess)
      ez = hipMemcpyAsync(P->zero_rows, zr.data(), sizeof(int) * zr.size(),
                          hipMemcpyHostToDevice, st);
    if (ez == hipSuccess) ez = hipStreamSynchronize(st);
    if (ez != hipSuccess) {
      oms_fold_destroy(P);
      SSLAB_HIP_CHECK(ez);
    }
  }
  *out = P;
  return kOk;
}

// X = M F for r columns (the backward sweep's entry state), with the maps in T.
template <typename T>
static int fold_apply_t(const FoldPlan* P, const T* maps, const T* F, int ldF, int r, T* X,
                        int ldx, hipStream_t st) {
  // long groups accumulate (split-K) into zeroed blocks; short ones store once, unsplit
  if (P->nzero > 0) {
    const long long n = (long long)P->nzero * P->m * r;
    const long long gsz = (n + 255) / 256;
    hipLaunchKernelGGL(k_zero_blocks<T>, dim3((unsigned)(gsz < 8192 ? gsz : 8192)), dim3(256), 0,
                       st, P->nzero, P->zero_rows, P->m, r, X, ldx);
    SSLAB_HIP_CHECK(hipGetLastError());
  }
  GemmLaunchT<T> G{};
  G.nd = 0;
  for (const FoldGroup& g : P->groups) {
    if (G.nd == kMaxDesc) {  // groups write disjoint blocks: any split into launches works
      SSLAB_TRY(gemm_launch(G, st));
      G.nd = 0;
    }
    T* C = X + (long long)g.x_row * ldx;
    const long long sC = (long long)g.x_rstride * ldx;
    int kt = 0;
    for (int t = 0; t < g.nterms; ++t) kt += g.t[t].K;
    const bool split_ok = kt >= P->nosplit_k;
    GemmDescT<T> d = split_ok ? gemm_desc<T>(P->m, r, g.batch, C, ldx, sC, C, ldx, sC, 1.0)
                              : gemm_desc<T>(P->m, r, g.batch, C, ldx, sC, nullptr, 0, 0, 0.0);
    d.nosplit = split_ok ? 0 : 1;
    for (int t = 0; t < g.nterms; ++t) {
      const FoldTerm& Tm = g.t[t];
      gemm_add_term<T>(d, gemm_term<T>(maps + Tm.a_off, Tm.lda, Tm.sA, 0,
                                       F + (long long)Tm.f_row * ldF, ldF,
                                       (long long)Tm.f_rstride * ldF, 0, Tm.K, 1.0));
    }
    G.d[G.nd++] = d;
  }
  if (G.nd > 0) SSLAB_TRY(gemm_launch(G, st));
  return kOk;
}

int oms_fold_apply(const FoldPlan* P, const double* F, int ldF, int r, double* X, int ldx,
                   hipStream_t st) {
  SSLAB_REQUIRE(P != nullptr, "oms_fold_apply: null plan");
  SSLAB_REQUIRE(r >= 0, "oms_fold_apply: negative column count");
  if (r == 0) return kOk;
  SSLAB_REQUIRE(F && X && ldF >= r && ldx >= r, "oms_fold_apply: bad pointers / strides");
  return fold_apply_t<double>(P, P->maps, F, ldF, r, X, ldx, st);
}

int oms_fold_enable_f32(FoldPlan* P, hipStream_t st) {
  SSLAB_REQUIRE(P != nullptr, "oms_fold_enable_f32: null plan");
  if (P->maps32 != nullptr) return kOk;
  if (P->maps_doubles > 0) {
    float* p = nullptr;
    SSLAB_HIP_CHECK(hipMalloc(&p, sizeof(float) * P->maps_doubles));
    const int rc = cast_f64_to_f32(P->maps_doubles, P->maps, p, st);
    if (rc != kOk) {
      (void)hipFree(p);
      return rc;
    }
    P->maps32 = p;
  }
  return kOk;
}

// Mixed-precision solve from boundary data: the FP32 maps and FP32 backward sweep give
// X0 (the folded solve in FP32), then `iters` refinement steps, each an FP64 residual
// r = B - S X, its FP32 red-black solve and an FP64 update.  B = rhs(F) in FP64 (the
// caller's), the system's diagonal blocks are identities (the oms S-system).
int oms_fold_solve_mixed(const FoldPlan* P, const float* factor32, const double* SiM,
                         const double* SiP, const double* F, int ldF, int r, const double* B,
                         double* X, double* work64, float* work32, int iters, hipStream_t st) {
  SSLAB_REQUIRE(P != nullptr, "oms_fold_solve_mixed: null plan");
  SSLAB_REQUIRE(r >= 0 && iters >= 0, "oms_fold_solve_mixed: bad sizes");
  if (r == 0) return kOk;
  SSLAB_REQUIRE(P->maps32 != nullptr || P->maps_doubles == 0,
                "oms_fold_solve_mixed: the plan has no FP32 maps (oms_fold_enable_f32)");
  SSLAB_REQUIRE(factor32 && F && B && X && work32 &&
                    (iters == 0 || (SiM && SiP && work64)) && (P->nb == 1 || (SiM && SiP)),
                "oms_fold_solve_mixed: null pointer");
  const long long n = (long long)P->nb * P->m * r;
  float* X32 = work32;
  float* V32 = X32 + n;
  float* F32 = V32 + n;
  SSLAB_REQUIRE(ldF == r, "oms_fold_solve_mixed: F must be contiguous (ldF == nrhs)");
  // F32 = F, then X32 = maps32 F32 and the FP32 backward sweep
  SSLAB_TRY(cast_f64_to_f32((long long)P->nF * r, F, F32, st));
  SSLAB_TRY(fold_apply_t<float>(P, P->maps32, F32, r, r, X32, r, st));
  SSLAB_TRY(rb_solve_f32_ex(P->nb, P->m, r, factor32, nullptr, r, X32, r, 1 | 8, st));
  SSLAB_TRY(cast_f32_to_f64(n, X32, X, st));
  return rb_refine_mixed(P->nb, P->m, r, factor32, nullptr, SiM, SiP, B, X, work64, V32, X32,
                         iters, st);
}

long long oms_fold_bytes(const FoldPlan* P) {
  if (!P) return 0;
  return (long long)sizeof(double) * P->maps_doubles +
         (P->maps32 ? (long long)sizeof(float) * P->maps_doubles : 0);
}

int oms_fold_groups(const FoldPlan* P) { return P ? (int)P->groups.size() : 0; }

void oms_fold_destroy(FoldPlan* P) {
  if (!P) return;
  if (P->maps) (void)hipFree(P->maps);
  if (P->maps32) (void)hipFree(P->maps32);
  if (P->zero_rows) (void)hipFree(P->zero_rows);
  delete P;
}

}  // namespace sslab
