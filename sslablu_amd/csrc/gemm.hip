// Grouped / batched GEMM for gfx950 on MFMA, FP64 and FP32.
//
// Tiling: a workgroup computes a BM x BN tile of C with WM x WN wavefronts (64
// lanes each); a wavefront owns a (BM/WM) x (BN/WN) sub-tile made of 16x16 MFMA
// fragments.  K is consumed in BK-wide slabs staged through LDS, k-major with a
// pad, and the next slab is prefetched into registers while the current one is
// multiplied.  FP64 launches with few output tiles but long K split K over several
// workgroups that accumulate with native global FP atomics (FP32 launches at most 4 parts).
//
// Operand lane layouts (one element per lane):
//   A (16x4):  lane l holds A[l % 16][l / 16];   B (4x16): lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, item v holds
//     v_mfma_f64_16x16x4_f64:  D[(l/16) + 4*v][l % 16]
//     v_mfma_f32_16x16x4_f32:  D[4*(l/16) + v][l % 16]
// (both measured on gfx950).
#include "gemm.hpp"

#include <type_traits>

namespace sslab {

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

template <int BM_, int BN_, int BK_, int WM_, int WN_>
struct GemmCfg {
  static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
  static constexpr int NT = WM * WN * 64;
  static constexpr int WTM = BM / WM, WTN = BN / WN;  // wavefront tile
  static constexpr int FM = WTM / 16, FN = WTN / 16;   // MFMA fragments per wave
  static constexpr int LDA_S = BM + 16;                // padded k-major LDS rows
  static constexpr int LDB_S = BN + 16;
  static constexpr int A_PER_T = BM * BK / NT;
  static constexpr int B_PER_T = BK * BN / NT;
  static_assert(BM % (WM * 16) == 0 && BN % (WN * 16) == 0, "tile/wave mismatch");
  static_assert(BK % 4 == 0, "BK must be a multiple of the MFMA K (4)");
  static_assert((BM * BK) % NT == 0 && (BK * BN) % NT == 0, "load split");
};

// Write one wave's accumulator fragments: added atomically (split K, atomic descriptors:
// C already holds beta Cin) or stored as C = acc + beta Cin.  Every Cin value of the
// wave is loaded before the first store: Cin may be C itself (in-place products), so the
// compiler cannot move a load above an earlier store, and value by value the epilogue
// would wait for one memory round trip per element.  Fragment (i, j) covers rows
// row_base + 16 i + Mfma::row(lane, v), columns col_base + 16 j + lane % 16.
template <typename T, int FM, int FN>
__device__ __forceinline__ void store_acc(const typename Mfma<T>::v4 (&acc)[FM][FN], T* C,
                                          int ldc, const T* Cin, int ldcin, T beta, bool split,
                                          int row_base, int col_base, int M, int N, int lane) {
  using M4 = Mfma<T>;
  const bool use_cin = !split && Cin != nullptr && beta != T(0);
  T cv[FM][FN][4];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = col_base + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = row_base + i * 16 + M4::row(lane, v);
        cv[i][j][v] = (use_cin && row < M && col < N) ? Cin[(long long)row * ldcin + col] : T(0);
      }
    }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = col_base + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = row_base + i * 16 + M4::row(lane, v);
        if (row < M && col < N) {
          if (split) atomicAdd(&C[(long long)row * ldc + col], acc[i][j][v]);
          else C[(long long)row * ldc + col] = acc[i][j][v] + beta * cv[i][j][v];
        }
      }
    }
}

// Map the hardware block id to a logical tile (the dispatcher assigns workgroups
// round-robin to the 8 XCDs, each with its own L2).  Chunks of cz consecutive logical
// tiles -- the tiles of one product, which share its A and B -- are dealt round-robin to
// the XCDs, so every XCD works on whole products and a launch whose products differ in
// cost (summed K, split parts) is spread evenly over the chip; the tail of fewer than
// 8 cz tiles is cut into 8 contiguous ranges.  Bijective on [0, n).
__device__ __forceinline__ int xcd_remap(int bid, int n, int cz) {
  const long long span = 8LL * cz;
  const int full = (int)(n / span * span);
  if (bid < full) {
    const int j = bid >> 3;  // position among the XCD's workgroups
    return ((bid & 7) + 8 * (j / cz)) * cz + j % cz;
  }
  const int b = bid - full, q = (n - full) >> 3;
  if (b < (q << 3)) return full + (b & 7) * q + (b >> 3);
  return bid;
}

template <class Cfg, typename T>
__device__ __forceinline__ void load_a(const GemmTermT<T>& t, const T* A, int M, int row0, int k0,
                                       int K, int tid, T (&ra)[Cfg::A_PER_T]) {
  if (!t.transA) {  // contiguous along k
#pragma unroll
    for (int r = 0; r < Cfg::A_PER_T; ++r) {
      const int e = tid + r * Cfg::NT;
      const int k = e % Cfg::BK, i = e / Cfg::BK;
      const int gi = row0 + i, gk = k0 + k;
      ra[r] = (gi < M && gk < K) ? A[(long long)gi * t.lda + gk] : T(0);
    }
  } else {
#pragma unroll
    for (int r = 0; r < Cfg::A_PER_T; ++r) {
      const int e = tid + r * Cfg::NT;
      const int i = e % Cfg::BM, k = e / Cfg::BM;
      const int gi = row0 + i, gk = k0 + k;
      ra[r] = (gi < M && gk < K) ? A[(long long)gk * t.lda + gi] : T(0);
    }
  }
}

template <class Cfg, typename T>
__device__ __forceinline__ void store_a(const GemmTermT<T>& t, int tid, T alpha,
                                        const T (&ra)[Cfg::A_PER_T], T* As) {
#pragma unroll
  for (int r = 0; r < Cfg::A_PER_T; ++r) {
    const int e = tid + r * Cfg::NT;
    int k, i;
    if (!t.transA) {
      k = e % Cfg::BK;
      i = e / Cfg::BK;
    } else {
      i = e % Cfg::BM;
      k = e / Cfg::BM;
    }
    As[k * Cfg::LDA_S + i] = alpha * ra[r];
  }
}

template <class Cfg, typename T>
__device__ __forceinline__ void load_b(const GemmTermT<T>& t, const T* B, int N, int col0, int k0,
                                       int K, int tid, T (&rb)[Cfg::B_PER_T]) {
  if (!t.transB) {  // contiguous along j
#pragma unroll
    for (int r = 0; r < Cfg::B_PER_T; ++r) {
      const int e = tid + r * Cfg::NT;
      const int j = e % Cfg::BN, k = e / Cfg::BN;
      const int gj = col0 + j, gk = k0 + k;
      rb[r] = (gj < N && gk < K) ? B[(long long)gk * t.ldb + gj] : T(0);
    }
  } else {
#pragma unroll
    for (int r = 0; r < Cfg::B_PER_T; ++r) {
      const int e = tid + r * Cfg::NT;
      const int k = e % Cfg::BK, j = e / Cfg::BK;
      const int gj = col0 + j, gk = k0 + k;
      rb[r] = (gj < N && gk < K) ? B[(long long)gj * t.ldb + gk] : T(0);
    }
  }
}

template <class Cfg, typename T>
__device__ __forceinline__ void store_b(const GemmTermT<T>& t, int tid, const T (&rb)[Cfg::B_PER_T],
                                        T* Bs) {
#pragma unroll
  for (int r = 0; r < Cfg::B_PER_T; ++r) {
    const int e = tid + r * Cfg::NT;
    int k, j;
    if (!t.transB) {
      j = e % Cfg::BN;
      k = e / Cfg::BN;
    } else {
      k = e % Cfg::BK;
      j = e / Cfg::BK;
    }
    Bs[k * Cfg::LDB_S + j] = rb[r];
  }
}

template <class Cfg, typename T>
__global__ void __launch_bounds__(Cfg::NT) k_gemm(const GemmLaunchT<T> L, int total_tiles) {
  using M4 = Mfma<T>;
  using v4 = typename M4::v4;
  __shared__ T As[Cfg::BK * Cfg::LDA_S];
  __shared__ T Bs[Cfg::BK * Cfg::LDB_S];

  const int tile = xcd_remap(blockIdx.x, total_tiles, L.xcd_chunk);

  // Select the descriptor with a fully unrolled scan (constant kernarg offsets).
  GemmDescT<T> D = L.d[0];
#pragma unroll
  for (int d = 1; d < kMaxDesc; ++d)
    if (d < L.nd && tile >= L.d[d].tile_begin) D = L.d[d];

  const int per_part = D.tiles_m * D.tiles_n;
  int local = tile - D.tile_begin;
  const int bk = local / per_part;  // (batch, k part) pair
  local -= bk * per_part;
  const int b = bk / D.ksplit;
  const int kp = bk - b * D.ksplit;
  const int kbeg = kp * D.kchunk, kfin = kbeg + D.kchunk;  // K range over the summed terms
  const int tm = local / D.tiles_n;
  const int tn = local - tm * D.tiles_n;
  const int row0 = tm * Cfg::BM, col0 = tn * Cfg::BN;
  const int M = D.M, N = D.N;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / Cfg::WN, wn = wave % Cfg::WN;
  const int wrow = wm * Cfg::WTM, wcol = wn * Cfg::WTN;

  v4 acc[Cfg::FM][Cfg::FN];
#pragma unroll
  for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
    for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = v4{T(0), T(0), T(0), T(0)};

  T ra[Cfg::A_PER_T], rb[Cfg::B_PER_T];

  int koff = 0;  // start of the current term in the summed K range
#pragma unroll
  for (int ti = 0; ti < kMaxTerms; ++ti) {
    if (ti >= D.nterms) break;
    const GemmTermT<T> Tm = D.term[ti];
    const T* A = Tm.A + (long long)b * Tm.strideA;
    const T* B = Tm.B + (long long)b * Tm.strideB;
    // this workgroup's part of the term: [ks, ke) in the term's own K coordinates
    const int ks = max(kbeg - koff, 0);
    const int ke = min(kfin - koff, Tm.K);
    koff += Tm.K;
    if (ke <= ks) continue;
    load_a<Cfg, T>(Tm, A, M, row0, ks, ke, tid, ra);
    load_b<Cfg, T>(Tm, B, N, col0, ks, ke, tid, rb);
    for (int k0 = ks; k0 < ke; k0 += Cfg::BK) {
      __syncthreads();
      store_a<Cfg, T>(Tm, tid, Tm.alpha, ra, As);
      store_b<Cfg, T>(Tm, tid, rb, Bs);
      __syncthreads();
      if (k0 + Cfg::BK < ke) {
        load_a<Cfg, T>(Tm, A, M, row0, k0 + Cfg::BK, ke, tid, ra);
        load_b<Cfg, T>(Tm, B, N, col0, k0 + Cfg::BK, ke, tid, rb);
      }
#pragma unroll
      for (int kk = 0; kk < Cfg::BK; kk += 4) {
        const int kr = kk + (lane >> 4);
        T af[Cfg::FM], bf[Cfg::FN];
#pragma unroll
        for (int i = 0; i < Cfg::FM; ++i) af[i] = As[kr * Cfg::LDA_S + wrow + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j) bf[j] = Bs[kr * Cfg::LDB_S + wcol + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
          for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = M4::op(af[i], bf[j], acc[i][j]);
      }
    }
  }

  T* C = D.C + (long long)b * D.strideC;
  const T* Cin = D.Cin ? D.Cin + (long long)b * D.strideCin : nullptr;
  store_acc<T, Cfg::FM, Cfg::FN>(acc, C, D.ldc, Cin, D.ldcin, D.beta, D.ksplit > 1 || D.atomic,
                                 row0 + wrow, col0 + wcol, M, N, lane);
}

// ---- fast path: every term as stored (row-major), 16-byte aligned operands ------------
//
// Same descriptors, split-K and epilogue as k_gemm, with
//   * 16-byte global loads; M and N edges clamp the row / column (the clamped values
//     only feed outputs that are never stored), K tails are zero-filled per vector;
//   * a two-stage LDS ring: one barrier per K slab, the next slab's global loads in
//     flight while the current slab is multiplied;
//   * both operands stored k-contiguous in LDS ([row][k] for A, [col][k] for B) and a
//     permuted k order: at MFMA step s, lane group g = lane / 16 supplies
//     k = g * BK/4 + s.  The MFMA sums over its four k lanes, so any assignment of k to
//     (step, lane group) shared by A and B gives the same product, and with this one a
//     lane's operands for consecutive steps are adjacent: one ds_read_b128 fetches
//     16 / sizeof(T) steps.
// The LDS row length BK + 16 / sizeof(T) makes the 16 row reads of a lane group hit
// disjoint banks.

template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};

template <class Cfg, typename T>
struct FastGeom {
  static constexpr int VEC = 16 / (int)sizeof(T);
  static constexpr int KS = Cfg::BK / 4;  // k values per lane group per slab
  static constexpr int LD = Cfg::BK + VEC;
  static constexpr int A_VECS = Cfg::BM * Cfg::BK / (VEC * Cfg::NT);
  static constexpr int B_VECS = Cfg::BK * Cfg::BN / (VEC * Cfg::NT);
  static constexpr int STAGE = (Cfg::BM + Cfg::BN) * LD;
  static constexpr bool ok = (Cfg::BM * Cfg::BK) % (VEC * Cfg::NT) == 0 &&
                             (Cfg::BK * Cfg::BN) % (VEC * Cfg::NT) == 0 && KS % VEC == 0 &&
                             A_VECS >= 1 && B_VECS >= 1 &&
                             2 * STAGE * (int)sizeof(T) <= 160 * 1024;  // LDS per CU
};

// The K loop of one term [ks, ke) for one workgroup, accumulating into acc.  VA / VB:
// 16-byte loads of A / B (chosen once per term; see k_gemm_fast).  Every load is
// unconditional -- out-of-range rows / columns / k read a clamped in-range address and
// the value is replaced by zero (k) or never stored (rows, columns) -- so the loads of
// a slab carry no branches, which would make the compiler wait for each load in turn.
template <class Cfg, typename T, bool VA, bool VB>
__device__ __forceinline__ void fast_term(const T* __restrict__ A, int lda,
                                          const T* __restrict__ B, int ldb, T alpha, int ks,
                                          int ke, int M, int N, int row0, int col0, int tid,
                                          int wrow, int wcol, int g, int li, T* smem, int& stage,
                                          typename Mfma<T>::v4 (&acc)[Cfg::FM][Cfg::FN]) {
  using G = FastGeom<Cfg, T>;
  using M4 = Mfma<T>;
  using vec = typename Vec16<T>::type;
  constexpr int VEC = G::VEC, KS = G::KS, LD = G::LD, BK = Cfg::BK;
  constexpr int KV = BK / VEC;       // vectors per A row slab
  constexpr int NV = Cfg::BN / VEC;  // vectors per B k row
  const int nslab = (ke - ks + BK - 1) / BK;
  vec ra[G::A_VECS], rb[G::B_VECS];

  // Raw loaded values stay in ra / rb while the current slab is multiplied; the
  // zero-fill of out-of-range k is applied when they are written to LDS, so nothing
  // consumes a load before the compute.
  bool oka[G::A_VECS], okb[G::B_VECS];
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < G::A_VECS; ++r) {
      const int e = tid + r * Cfg::NT;
      const int kv = e % KV, i = e / KV;
      const T* src = A + (long long)min(row0 + i, M - 1) * lda;
      const int k = k0 + kv * VEC;
      if (VA) {  // [ks, ke) holds whole vectors: a vector is all in or all out
        ra[r] = *reinterpret_cast<const vec*>(src + min(k, ke - VEC));
        oka[r] = k < ke;
      } else {
#pragma unroll
        for (int q = 0; q < VEC; ++q) ra[r][q] = src[min(k + q, ke - 1)];
        oka[r] = true;  // per-element k checks at the store
      }
    }
#pragma unroll
    for (int r = 0; r < G::B_VECS; ++r) {
      const int e = tid + r * Cfg::NT;
      const int jv = e % NV, kk = e / NV;
      const int k = k0 + kk;
      const T* src = B + (long long)min(k, ke - 1) * ldb;
      const int c = col0 + jv * VEC;
      if (VB) {  // N is a multiple of VEC: a vector is all in or all out of [0, N)
        rb[r] = *reinterpret_cast<const vec*>(src + min(c, N - VEC));
      } else {
#pragma unroll
        for (int q = 0; q < VEC; ++q) rb[r][q] = src[min(c + q, N - 1)];
      }
      okb[r] = k < ke;
    }
  };
  auto store = [&](int st, int k0) {
    T* As = smem + st * G::STAGE;
    T* Bs = As + Cfg::BM * LD;
#pragma unroll
    for (int r = 0; r < G::A_VECS; ++r) {
      const int e = tid + r * Cfg::NT;
      const int kv = e % KV, i = e / KV;
      vec v = alpha * ra[r];
      if (VA) {
        v = oka[r] ? v : vec(T(0));
      } else {
        const int k = k0 + kv * VEC;
#pragma unroll
        for (int q = 0; q < VEC; ++q) v[q] = (k + q < ke) ? v[q] : T(0);
      }
      *reinterpret_cast<vec*>(As + i * LD + kv * VEC) = v;
    }
#pragma unroll
    for (int r = 0; r < G::B_VECS; ++r) {
      const int e = tid + r * Cfg::NT;
      const int jv = e % NV, kk = e / NV;
#pragma unroll
      for (int q = 0; q < VEC; ++q) Bs[(jv * VEC + q) * LD + kk] = okb[r] ? rb[r][q] : T(0);
    }
  };

  load(ks);
  store(stage, ks);
  __syncthreads();
  for (int t = 0; t < nslab; ++t) {
    const bool more = t + 1 < nslab;
    if (more) load(ks + (t + 1) * BK);
    {
      const T* As = smem + stage * G::STAGE;
      const T* Bs = As + Cfg::BM * LD;
      // one 16-byte group of k steps at a time keeps the operand registers small
#pragma unroll
      for (int q = 0; q < KS / VEC; ++q) {
        vec af[Cfg::FM], bf[Cfg::FN];
#pragma unroll
        for (int i = 0; i < Cfg::FM; ++i)
          af[i] = *reinterpret_cast<const vec*>(As + (wrow + i * 16 + li) * LD + g * KS + q * VEC);
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j)
          bf[j] = *reinterpret_cast<const vec*>(Bs + (wcol + j * 16 + li) * LD + g * KS + q * VEC);
#pragma unroll
        for (int e = 0; e < VEC; ++e)
#pragma unroll
          for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
            for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = M4::op(af[i][e], bf[j][e], acc[i][j]);
      }
    }
    if (more) store(stage ^ 1, ks + (t + 1) * BK);
    __syncthreads();
    stage ^= 1;
  }
}

// Aligned operands (every term takes 16-byte loads): one slab sequence over all terms
// of the workgroup's K range, with the global loads of two slabs in flight (register
// sets 0 and 1 alternate) while a third is multiplied from LDS.  A slab therefore waits
// for loads issued two compute phases earlier, and a term boundary does not drain the
// pipeline.  Loads past the last slab re-read in-range vectors and are never stored to a
// stage that is multiplied.
template <class Cfg, typename T>
__device__ __forceinline__ void fast_desc_pipe(const GemmDescT<T>& D, int b, int kbeg, int kfin,
                                               int M, int N, int row0, int col0, int tid,
                                               int wrow, int wcol, int g, int li, T* smem,
                                               typename Mfma<T>::v4 (&acc)[Cfg::FM][Cfg::FN]) {
  using G = FastGeom<Cfg, T>;
  using M4 = Mfma<T>;
  using vec = typename Vec16<T>::type;
  constexpr int VEC = G::VEC, KS = G::KS, LD = G::LD, BK = Cfg::BK;
  constexpr int KV = BK / VEC;       // vectors per A row slab
  constexpr int NV = Cfg::BN / VEC;  // vectors per B k row
  const int nterms = D.nterms;

  int S = 0;  // slabs over all terms
  {
    int off = 0;
    for (int ti = 0; ti < nterms; ++ti) {
      const int K = D.term[ti].K;
      const int ks = max(kbeg - off, 0), ke = min(kfin - off, K);
      off += K;
      if (ke > ks) S += (ke - ks + BK - 1) / BK;
    }
  }
  if (S == 0) return;

  // load cursor: operands of the current term, next k to load and the term's k end
  int cti = -1, cstart = 0, ck = 0, cke = 0, clda = 0, cldb = 0;
  const T* cA = nullptr;
  const T* cB = nullptr;
  T calpha = T(0);
  auto next_term = [&]() {
    for (int ti = cti + 1; ti < nterms; ++ti) {
      const GemmTermT<T>& Tm = D.term[ti];
      const int off = cstart;
      cstart += Tm.K;
      const int ks = max(kbeg - off, 0), ke = min(kfin - off, Tm.K);
      if (ke > ks) {
        cti = ti;
        cA = Tm.A + (long long)b * Tm.strideA;
        cB = Tm.B + (long long)b * Tm.strideB;
        clda = Tm.lda;
        cldb = Tm.ldb;
        ck = ks;
        cke = ke;
        calpha = Tm.alpha;
        return;
      }
    }
    cti = nterms;  // exhausted: k runs on past cke, loads stay clamped in range
  };

  vec ra[2][G::A_VECS], rb[2][G::B_VECS];
  bool oka[2][G::A_VECS], okb[2][G::B_VECS];
  T al[2];
  auto load = [&](auto set) {
    constexpr int s = decltype(set)::value;
    al[s] = calpha;
#pragma unroll
    for (int r = 0; r < G::A_VECS; ++r) {
      const int e = tid + r * Cfg::NT;
      const int kv = e % KV, i = e / KV;
      const T* src = cA + (long long)min(row0 + i, M - 1) * clda;
      const int k = ck + kv * VEC;
      ra[s][r] = *reinterpret_cast<const vec*>(src + min(k, cke - VEC));
      oka[s][r] = k < cke;
    }
#pragma unroll
    for (int r = 0; r < G::B_VECS; ++r) {
      const int e = tid + r * Cfg::NT;
      const int jv = e % NV, kk = e / NV;
      const int k = ck + kk;
      const T* src = cB + (long long)min(k, cke - 1) * cldb;
      rb[s][r] = *reinterpret_cast<const vec*>(src + min(col0 + jv * VEC, N - VEC));
      okb[s][r] = k < cke;
    }
    ck += BK;
    if (ck >= cke && cti < nterms) next_term();
  };
  auto store = [&](auto set) {
    constexpr int s = decltype(set)::value;
    T* As = smem + s * G::STAGE;
    T* Bs = As + Cfg::BM * LD;
#pragma unroll
    for (int r = 0; r < G::A_VECS; ++r) {
      const int e = tid + r * Cfg::NT;
      const int kv = e % KV, i = e / KV;
      const vec v = al[s] * ra[s][r];
      *reinterpret_cast<vec*>(As + i * LD + kv * VEC) = oka[s][r] ? v : vec(T(0));
    }
#pragma unroll
    for (int r = 0; r < G::B_VECS; ++r) {
      const int e = tid + r * Cfg::NT;
      const int jv = e % NV, kk = e / NV;
#pragma unroll
      for (int q = 0; q < VEC; ++q) Bs[(jv * VEC + q) * LD + kk] = okb[s][r] ? rb[s][r][q] : T(0);
    }
  };
  auto compute = [&](int st) {
    const T* As = smem + st * G::STAGE;
    const T* Bs = As + Cfg::BM * LD;
#pragma unroll
    for (int q = 0; q < KS / VEC; ++q) {
      vec af[Cfg::FM], bf[Cfg::FN];
#pragma unroll
      for (int i = 0; i < Cfg::FM; ++i)
        af[i] = *reinterpret_cast<const vec*>(As + (wrow + i * 16 + li) * LD + g * KS + q * VEC);
#pragma unroll
      for (int j = 0; j < Cfg::FN; ++j)
        bf[j] = *reinterpret_cast<const vec*>(Bs + (wcol + j * 16 + li) * LD + g * KS + q * VEC);
#pragma unroll
      for (int e = 0; e < VEC; ++e)
#pragma unroll
        for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
          for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = M4::op(af[i][e], bf[j][e], acc[i][j]);
    }
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;

  next_term();
  load(I0{});  // slab 0
  load(I1{});  // slab 1
  store(I0{});
  __syncthreads();
  for (int t = 0; t < S; t += 2) {
    load(I0{});  // slab t+2 (register set 0 was stored to stage 0 before)
    compute(0);  // slab t
    store(I1{});  // slab t+1 into stage 1 (last read two phases ago)
    __syncthreads();
    if (t + 1 >= S) break;
    load(I1{});  // slab t+3
    compute(1);  // slab t+1
    store(I0{});  // slab t+2
    __syncthreads();
  }
}

// Two waves per SIMD at least (256 registers per lane): the 128x128 tile otherwise
// takes 268 and runs one wave per SIMD.
// kAligned: the host has checked that every term takes 16-byte loads for A and B (then
// only that path is compiled, which keeps the register count of the common case low).
// One output tile of k_gemm_fast: tile index `tile` of the launch, which belongs to
// descriptor D (D.tile_begin <= tile).
template <class Cfg, typename T, int kMode>
__device__ __forceinline__ void fast_tile(const GemmDescT<T>& D, int tile, T* smem) {
  constexpr bool kAligned = kMode >= 1;
  using G = FastGeom<Cfg, T>;
  using M4 = Mfma<T>;
  using v4 = typename M4::v4;
  constexpr int VEC = G::VEC;

  const int per_part = D.tiles_m * D.tiles_n;
  int local = tile - D.tile_begin;
  const int bk = local / per_part;
  local -= bk * per_part;
  const int b = bk / D.ksplit;
  const int kp = bk - b * D.ksplit;
  const int kbeg = kp * D.kchunk, kfin = kbeg + D.kchunk;
  const int tm = local / D.tiles_n;
  const int tn = local - tm * D.tiles_n;
  const int row0 = tm * Cfg::BM, col0 = tn * Cfg::BN;
  const int M = D.M, N = D.N;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / Cfg::WN, wn = wave % Cfg::WN;
  const int wrow = wm * Cfg::WTM, wcol = wn * Cfg::WTN;
  const int g = lane >> 4, li = lane & 15;

  v4 acc[Cfg::FM][Cfg::FN];
#pragma unroll
  for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
    for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = v4{T(0), T(0), T(0), T(0)};

  int stage = 0;
  int koff = 0;
  const int nterms = kMode == 2 ? 0 : D.nterms;
  if constexpr (kMode == 2)
    fast_desc_pipe<Cfg, T>(D, b, kbeg, kfin, M, N, row0, col0, tid, wrow, wcol, g, li, smem, acc);
  for (int ti = 0; ti < nterms; ++ti) {
    const GemmTermT<T>& Tm = D.term[ti];
    const int ks = max(kbeg - koff, 0);
    const int ke = min(kfin - koff, Tm.K);
    koff += Tm.K;
    if (ke <= ks) continue;
    const T* A = Tm.A + (long long)b * Tm.strideA;
    const T* B = Tm.B + (long long)b * Tm.strideB;
    // 16-byte loads need aligned rows; for A also K-range ends on vector boundaries,
    // for B a column count that is a multiple of the vector (uniform per term).
    const bool va = ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && (Tm.lda % VEC == 0) &&
                    (ks % VEC == 0) && (ke % VEC == 0);
    const bool vb = ((reinterpret_cast<uintptr_t>(B) & 15) == 0) && (Tm.ldb % VEC == 0) &&
                    (N % VEC == 0);
    if (kAligned || (va && vb))
      fast_term<Cfg, T, true, true>(A, Tm.lda, B, Tm.ldb, Tm.alpha, ks, ke, M, N, row0, col0, tid,
                                    wrow, wcol, g, li, smem, stage, acc);
    else if (kAligned) {
    } else if (va)
      fast_term<Cfg, T, true, false>(A, Tm.lda, B, Tm.ldb, Tm.alpha, ks, ke, M, N, row0, col0, tid,
                                     wrow, wcol, g, li, smem, stage, acc);
    else if (vb)
      fast_term<Cfg, T, false, true>(A, Tm.lda, B, Tm.ldb, Tm.alpha, ks, ke, M, N, row0, col0, tid,
                                     wrow, wcol, g, li, smem, stage, acc);
    else
      fast_term<Cfg, T, false, false>(A, Tm.lda, B, Tm.ldb, Tm.alpha, ks, ke, M, N, row0, col0,
                                      tid, wrow, wcol, g, li, smem, stage, acc);
  }

  T* C = D.C + (long long)b * D.strideC;
  const T* Cin = D.Cin ? D.Cin + (long long)b * D.strideCin : nullptr;
  store_acc<T, Cfg::FM, Cfg::FN>(acc, C, D.ldc, Cin, D.ldcin, D.beta, D.ksplit > 1 || D.atomic,
                           row0 + wrow, col0 + wcol, M, N, lane);
}

template <class Cfg, typename T, int kMode>
__global__ void __launch_bounds__(Cfg::NT) __attribute__((amdgpu_waves_per_eu(2)))
k_gemm_fast(const GemmLaunchT<T> L, int total_tiles) {
  __shared__ __attribute__((aligned(16))) T smem[2 * FastGeom<Cfg, T>::STAGE];
  const int tile = xcd_remap(blockIdx.x, total_tiles, L.xcd_chunk);
  int di = 0;
#pragma unroll
  for (int d = 1; d < kMaxDesc; ++d)
    if (d < L.nd && tile >= L.d[d].tile_begin) di = d;
  fast_tile<Cfg, T, kMode>(L.d[di], tile, smem);
}

// ---- dependency-ordered launches -------------------------------------------------------
//
// k_gemm_chain runs the tiles of several dependent product levels in one launch.  Tiles
// are taken in launch order from a ticket counter (L.ctl[0]), so when a workgroup holds
// tile t every tile below t is held by a workgroup that is already running.  A tile
// first waits until, for each block its terms read (dep1 / dep_stride), the counter of
// that block has reached the number of this launch's tiles that write it (need[]); the
// writers all come before it in launch order, so the wait always ends.  After its
// stores and atomics it adds one to the counter of the block it writes (sig1 /
// sig_stride).  Waits are bounded: past spin_limit polls the workgroup sets the abort
// flag L.ctl[2] and goes on (its result is then wrong and the host reports it).  The
// last workgroup to finish clears the ticket, the finish count and every block counter,
// so the next launch starts from zero without a fill kernel.
//
// Memory order: every wave of the writer issues a device-scope release fence (waits for
// its own stores and atomics, writes the XCD's L2 back) before the workgroup barrier
// after which one thread increments the counter; the reader polls with relaxed
// device-scope loads and every wave issues one device-scope acquire fence (L1 / L2
// invalidation) after the wait.
__device__ __forceinline__ int chain_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class Cfg, typename T, int kMode>
__global__ void __launch_bounds__(Cfg::NT) __attribute__((amdgpu_waves_per_eu(2)))
k_gemm_chain(const GemmChainT<T> L, int total_tiles) {
  __shared__ __attribute__((aligned(16))) T smem[2 * FastGeom<Cfg, T>::STAGE];
  __shared__ int s_tile;
  const int tid = threadIdx.x;
  if (tid == 0) s_tile = __hip_atomic_fetch_add(&L.ctl[0], 1, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int tile = s_tile;
  if (tile < 0 || tile >= total_tiles) {  // counters not at zero on entry: give up
    if (tid == 0) {
      __hip_atomic_store(&L.ctl[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int done = __hip_atomic_fetch_add(&L.ctl[1], 1, __ATOMIC_ACQ_REL,
                                              __HIP_MEMORY_SCOPE_AGENT);
      if (done == total_tiles - 1) {
        for (int i = 0; i < kChainBlocks; ++i)
          __hip_atomic_store(&L.cnt[i], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&L.ctl[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&L.ctl[1], 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    return;
  }
  int di = 0;
#pragma unroll
  for (int d = 1; d < kChainDesc; ++d)
    if (d < L.nd && tile >= L.d[d].tile_begin) di = d;
  const GemmDescT<T>& D = L.d[di];
  const int b = (tile - D.tile_begin) / (D.tiles_m * D.tiles_n * D.ksplit);
  if (tid == 0) {
    for (int t = 0; t < D.nterms; ++t) {
      if (L.dep1[di][t] == 0) continue;
      const int blk = L.dep1[di][t] - 1 + b * L.dep_stride[di][t];
      const int need = L.need[blk];
      int it = 0;
      while (chain_load(&L.cnt[blk]) < need) {
        if (++it > L.spin_limit || chain_load(&L.ctl[2]) != 0) {
          __hip_atomic_store(&L.ctl[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // drop stale L1 / L2 lines
  fast_tile<Cfg, T, kMode>(D, tile, smem);
  // every wave waits for its own stores / atomics and writes the XCD's L2 back, then one
  // thread signals
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (tid == 0) {
    if (L.sig1[di] != 0) {
      __hip_atomic_fetch_add(&L.cnt[L.sig1[di] - 1 + b * L.sig_stride[di]], 1,
                             __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    const int done = __hip_atomic_fetch_add(&L.ctl[1], 1, __ATOMIC_ACQ_REL,
                                            __HIP_MEMORY_SCOPE_AGENT);
    if (done == total_tiles - 1) {  // every other tile has finished: reset for the next launch
      for (int i = 0; i < kChainBlocks; ++i)
        __hip_atomic_store(&L.cnt[i], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&L.ctl[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&L.ctl[1], 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}


// ---- narrow N: A streamed from HBM straight into MFMA operand registers ----------------
//
// For few right-hand sides (N <= 64) every A element is used by exactly one wavefront,
// so staging A in LDS only costs LDS space, which caps the bytes in flight per CU.
// Here the wavefronts of a workgroup are stacked along M (WN = 1) and each loads its
// own A fragments with 16-byte loads in the permuted-k order of k_gemm_fast (lane group
// g holds k = g*BK/4 + s), one slab ahead in a second register buffer; only the B slab
// (shared by all wavefronts) goes through a two-stage LDS ring, pre-scaled by alpha.
// Aligned operands only (k ranges on vector boundaries); out-of-range k re-reads the
// last in-range A vector and meets zero rows of B.
template <class Cfg, typename T>
struct ADGeom {
  static constexpr int VEC = 16 / (int)sizeof(T);
  static constexpr int KS = Cfg::BK / 4;
  static constexpr int QV = KS / VEC;
  static constexpr int LD = Cfg::BK + VEC;
  static constexpr int NV = Cfg::BN / VEC;
  static constexpr int B_VECS = Cfg::BK * Cfg::BN / (VEC * Cfg::NT);
  static constexpr int STAGE = Cfg::BN * LD;
  static constexpr bool ok = Cfg::WN == 1 && (Cfg::BK * Cfg::BN) % (VEC * Cfg::NT) == 0 &&
                             KS % VEC == 0 && B_VECS >= 1;
};

template <class Cfg, typename T>
__global__ void __launch_bounds__(Cfg::NT) __attribute__((amdgpu_waves_per_eu(2)))
k_gemm_adirect(const GemmLaunchT<T> L, int total_tiles) {
  using G = ADGeom<Cfg, T>;
  using M4 = Mfma<T>;
  using v4 = typename M4::v4;
  using vec = typename Vec16<T>::type;
  constexpr int VEC = G::VEC, KS = G::KS, QV = G::QV, LD = G::LD, NV = G::NV, BK = Cfg::BK;
  constexpr int FM = Cfg::FM, FN = Cfg::FN;
  __shared__ __attribute__((aligned(16))) T smem[2 * G::STAGE];

  const int tile = xcd_remap(blockIdx.x, total_tiles, L.xcd_chunk);
  int di = 0;
#pragma unroll
  for (int d = 1; d < kMaxDesc; ++d)
    if (d < L.nd && tile >= L.d[d].tile_begin) di = d;
  const GemmDescT<T>& D = L.d[di];
  const int per_part = D.tiles_m * D.tiles_n;
  int local = tile - D.tile_begin;
  const int bk = local / per_part;
  local -= bk * per_part;
  const int b = bk / D.ksplit;
  const int kp = bk - b * D.ksplit;
  const int kbeg = kp * D.kchunk, kfin = kbeg + D.kchunk;
  const int tm = local / D.tiles_n;
  const int tn = local - tm * D.tiles_n;
  const int row0 = tm * Cfg::BM, col0 = tn * Cfg::BN;
  const int M = D.M, N = D.N;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wrow = (tid >> 6) * Cfg::WTM;
  const int g = lane >> 4, li = lane & 15;

  v4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4{T(0), T(0), T(0), T(0)};

  int stage = 0;
  int koff = 0;
  const int nterms = D.nterms;
  for (int ti = 0; ti < nterms; ++ti) {
    const GemmTermT<T>& Tm = D.term[ti];
    const int ks = max(kbeg - koff, 0);
    const int ke = min(kfin - koff, Tm.K);
    koff += Tm.K;
    if (ke <= ks) continue;
    const T* A = Tm.A + (long long)b * Tm.strideA;
    const T* B = Tm.B + (long long)b * Tm.strideB;
    const int lda = Tm.lda, ldb = Tm.ldb;
    const T alpha = Tm.alpha;
    const int nslab = (ke - ks + BK - 1) / BK;
    const T* arow[FM];
#pragma unroll
    for (int fi = 0; fi < FM; ++fi)
      arow[fi] = A + (long long)min(row0 + wrow + fi * 16 + li, M - 1) * lda + g * KS;
    vec ra0[FM][QV], ra1[FM][QV], rb[G::B_VECS];
    bool okb[G::B_VECS];

    auto load_a = [&](int k0, vec (&ra)[FM][QV]) {
#pragma unroll
      for (int fi = 0; fi < FM; ++fi)
#pragma unroll
        for (int q = 0; q < QV; ++q)
          ra[fi][q] = *reinterpret_cast<const vec*>(
              arow[fi] + min(k0 + q * VEC, ke - VEC - g * KS));
    };
    auto load_b = [&](int k0) {
#pragma unroll
      for (int r = 0; r < G::B_VECS; ++r) {
        const int e = tid + r * Cfg::NT;
        const int jv = e % NV, kk = e / NV;
        const int k = k0 + kk;
        rb[r] = *reinterpret_cast<const vec*>(B + (long long)min(k, ke - 1) * ldb +
                                              min(col0 + jv * VEC, N - VEC));
        okb[r] = k < ke;
      }
    };
    auto store_b = [&](int st) {
      T* Bs = smem + st * G::STAGE;
#pragma unroll
      for (int r = 0; r < G::B_VECS; ++r) {
        const int e = tid + r * Cfg::NT;
        const int jv = e % NV, kk = e / NV;
        const vec v = okb[r] ? alpha * rb[r] : vec(T(0));
#pragma unroll
        for (int q = 0; q < VEC; ++q) Bs[(jv * VEC + q) * LD + kk] = v[q];
      }
    };
    auto compute = [&](const vec (&ra)[FM][QV], int st) {
      const T* Bs = smem + st * G::STAGE;
#pragma unroll
      for (int q = 0; q < QV; ++q) {
        vec bf[FN];
#pragma unroll
        for (int j = 0; j < FN; ++j)
          bf[j] = *reinterpret_cast<const vec*>(Bs + (j * 16 + li) * LD + g * KS + q * VEC);
#pragma unroll
        for (int e = 0; e < VEC; ++e)
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = M4::op(ra[i][q][e], bf[j][e], acc[i][j]);
      }
    };

    // The A address clamp min(k, ke - VEC) is taken relative to the lane's own offset
    // g*KS (folded into arow): lane position g*KS + k0 + q*VEC <= ke - VEC.
    load_a(ks, ra0);
    load_b(ks);
    store_b(stage);
    __syncthreads();
    for (int t = 0; t < nslab; t += 2) {
      const bool m1 = t + 1 < nslab;
      if (m1) {
        load_a(ks + (t + 1) * BK, ra1);
        load_b(ks + (t + 1) * BK);
      }
      compute(ra0, stage);
      if (m1) store_b(stage ^ 1);
      __syncthreads();
      stage ^= 1;
      if (!m1) break;
      const bool m2 = t + 2 < nslab;
      if (m2) {
        load_a(ks + (t + 2) * BK, ra0);
        load_b(ks + (t + 2) * BK);
      }
      compute(ra1, stage);
      if (m2) store_b(stage ^ 1);
      __syncthreads();
      stage ^= 1;
    }
  }

  T* C = D.C + (long long)b * D.strideC;
  const T* Cin = D.Cin ? D.Cin + (long long)b * D.strideCin : nullptr;
  store_acc<T, FM, FN>(acc, C, D.ldc, Cin, D.ldcin, D.beta, D.ksplit > 1 || D.atomic,
                           row0 + wrow, col0, M, N, lane);
}

// 0: some term uses a transposed operand (the LDS-transposing kernel k_gemm);
// 1: the vector-load kernel, per-term choice of 16-byte or element loads;
// 2: the vector-load kernel with 16-byte loads everywhere (aligned operands, leading
//    dimensions, every K and N multiples of the vector width).
template <typename T, class LaunchT>
static int fast_kind(const LaunchT& P) {
  constexpr int VEC = 16 / (int)sizeof(T);
  static const bool enabled = [] {
    const char* s = getenv("SSLAB_GEMM_FAST");
    return !(s && s[0] == '0');
  }();
  if (!enabled) return 0;
  bool aligned = true;
  for (int i = 0; i < P.nd; ++i) {
    const GemmDescT<T>& d = P.d[i];
    if (d.N % VEC) aligned = false;
    for (int t = 0; t < d.nterms; ++t) {
      const GemmTermT<T>& m = d.term[t];
      if (m.transA || m.transB) return 0;
      if ((reinterpret_cast<uintptr_t>(m.A) | reinterpret_cast<uintptr_t>(m.B)) & 15) aligned = false;
      if (m.lda % VEC || m.ldb % VEC || m.K % VEC || m.strideA % VEC || m.strideB % VEC)
        aligned = false;
    }
  }
  return aligned ? 2 : 1;
}

// Tile configurations, by id (the ids of GemmPolicy and sslab_gemm_set_policy):
//   0: 128x128 BK16   1: 128x64 BK16   2: 64x64 BK32   3: 64x64 BK16
//   4: 64x32 BK16     5: 32x32 BK16    6: 32x64 BK16 (2 waves)   7: 64x128 BK16
//   8: 32x16 BK32 (2 waves)   9: 32x32 BK32 (2 waves)     -- few right-hand sides
//  10: 128x64 BK32, 11: 128x64 BK64 -- A streamed to registers (k_gemm_adirect), aligned
//      launches with at most 64 columns; otherwise the same tile on k_gemm_fast
//  12: 32x32 BK32 (4 waves): twice the MFMA work of tile 5 per LDS stage and barrier
//  13: 32x32 BK64 (4 waves), 14: 64x32 BK32 (4 waves)
using Cfg128 = GemmCfg<128, 128, 16, 2, 2>;
using Cfg128x64 = GemmCfg<128, 64, 16, 2, 2>;
using Cfg64 = GemmCfg<64, 64, 32, 2, 2>;
using Cfg64s = GemmCfg<64, 64, 16, 2, 2>;
using Cfg64x32 = GemmCfg<64, 32, 16, 2, 2>;
using Cfg32 = GemmCfg<32, 32, 16, 2, 2>;
using Cfg32x64 = GemmCfg<32, 64, 16, 1, 2>;
using Cfg64x128 = GemmCfg<64, 128, 16, 2, 2>;
using Cfg32x16 = GemmCfg<32, 16, 32, 2, 1>;
using Cfg32w = GemmCfg<32, 32, 32, 2, 1>;
using Cfg32k = GemmCfg<32, 32, 32, 2, 2>;
using Cfg32k64 = GemmCfg<32, 32, 64, 2, 2>;
using Cfg64x32k = GemmCfg<64, 32, 32, 2, 2>;
using CfgAD32 = GemmCfg<128, 64, 32, 4, 1>;  // A in registers (k_gemm_adirect)
using CfgAD64 = GemmCfg<128, 64, 64, 4, 1>;
constexpr int kNumCfg = 15;
constexpr int kCfgBM[kNumCfg] = {128, 128, 64, 64, 64, 32, 32, 64, 32, 32, 128, 128, 32, 32, 64};
constexpr int kCfgBN[kNumCfg] = {128, 64, 64, 64, 32, 32, 64, 128, 16, 32, 64, 64, 32, 32, 32};

// Workgroups a launch would create with BM x BN tiles.
template <typename T>
static long long count_tiles(const GemmLaunchT<T>& L, int BM, int BN) {
  long long t = 0;
  for (int i = 0; i < L.nd; ++i) {
    const GemmDescT<T>& d = L.d[i];
    if (d.batch > 0 && d.M > 0 && d.N > 0)
      t += (long long)ceil_div(d.M, BM) * ceil_div(d.N, BN) * d.batch;
  }
  return t;
}

// The initial values of a launch's split-K outputs that do not accumulate in place,
// C_b = beta * Cin_b (Cin may be null or alias C), which the K parts then add into
// atomically: one table of outputs per GEMM launch and one init launch for all of them.
// Rows of all outputs are numbered in turn (output q owns rows row_begin.. of its batch x
// M); a workgroup writes one row at a time, 256 columns per pass.
template <typename T>
struct InitJob {
  T* C;
  const T* Cin;
  long long sC, sCin;
  int ldc, ldcin, M, N;
  T beta;
  int row_begin;
};
template <typename T>
struct InitTable {
  InitJob<T> j[kMaxDesc];
  int n, rows;
};

template <typename T>
__global__ __launch_bounds__(256) void k_scale_init(const InitTable<T> t) {
  for (int r = blockIdx.x; r < t.rows; r += gridDim.x) {
    InitJob<T> J = t.j[0];  // selected by an unrolled scan (constant kernarg offsets)
#pragma unroll
    for (int q = 1; q < kMaxDesc; ++q)
      if (q < t.n && r >= t.j[q].row_begin) J = t.j[q];
    const int lr = r - J.row_begin;
    const int b = lr / J.M, i = lr - b * J.M;
    T* c = J.C + b * J.sC + (long long)i * J.ldc;
    const bool use = J.Cin != nullptr && J.beta != T(0);
    const T* ci = J.Cin + (use ? b * J.sCin + (long long)i * J.ldcin : 0);
    for (int col = threadIdx.x; col < J.N; col += 256) c[col] = use ? J.beta * ci[col] : T(0);
  }
}

template <typename T>
static long long total_k(const GemmDescT<T>& d) {
  long long k = 0;
  for (int t = 0; t < d.nterms; ++t) k += d.term[t].K;
  return k;
}

// Split-K chunk of a launch with BM x BN tiles that aims at ksplit x (output tiles)
// workgroups: the shortest K range c (a multiple of BK, at least two BK slabs) for which
// splitting every descriptor into min(ceil(K_d / c), maxparts) parts stays within that
// budget.  Descriptors of equal K are then split ksplit ways; in a launch of mixed K the
// long products get more parts than the short ones, so no workgroup runs much longer than
// the rest.  Returns 0 when nothing is worth splitting.
template <class Cfg, typename T, class LaunchT>
static long long split_chunk(const LaunchT& L, int ksplit, int maxparts) {
  long long tiles = 0, kmax = 0;
  for (int i = 0; i < L.nd; ++i) {
    const GemmDescT<T>& d = L.d[i];
    if (d.batch <= 0 || d.M <= 0 || d.N <= 0) continue;
    tiles += (long long)ceil_div(d.M, Cfg::BM) * ceil_div(d.N, Cfg::BN) * d.batch;
    const long long K = total_k(d);
    kmax = K > kmax ? K : kmax;
  }
  const long long budget = tiles * ksplit;
  auto wgs = [&](long long c) {
    long long w = 0;
    for (int i = 0; i < L.nd; ++i) {
      const GemmDescT<T>& d = L.d[i];
      if (d.batch <= 0 || d.M <= 0 || d.N <= 0) continue;
      long long parts = d.nosplit ? 1 : (total_k(d) + c - 1) / c;
      parts = parts < 1 ? 1 : (parts > maxparts ? maxparts : parts);
      w += (long long)ceil_div(d.M, Cfg::BM) * ceil_div(d.N, Cfg::BN) * d.batch * parts;
    }
    return w;
  };
  long long lo = 2, hi = (kmax + Cfg::BK - 1) / Cfg::BK;  // chunk in BK slabs
  if (ksplit <= 1 || hi <= lo) return 0;
  while (lo < hi) {  // the workgroup count never grows with the chunk
    const long long mid = (lo + hi) / 2;
    if (wgs(mid * Cfg::BK) <= budget) hi = mid;
    else lo = mid + 1;
  }
  return lo * Cfg::BK;
}

// Tiles per XCD chunk (xcd_remap): the fewest tiles of one product (output tiles x K
// parts) over the launch's descriptors, so no XCD holds a run of one descriptor's products.
template <typename T>
static int xcd_chunk(const GemmLaunchT<T>& P, int total) {
  int cz = total;
  for (int i = 0; i < P.nd; ++i) {
    const int per = P.d[i].tiles_m * P.d[i].tiles_n * P.d[i].ksplit;
    cz = per < cz ? per : cz;
  }
  return cz < 1 ? 1 : cz;
}

// ksplit > 1 splits the output tiles' summed K ranges over about ksplit x (output tiles)
// workgroups in all (split_chunk), at most maxparts per tile.
template <class Cfg, typename T, bool kADirect = false>
static int launch_cfg(GemmLaunchT<T>& L, hipStream_t stream, int ksplit, int maxparts = 64) {
  int total = 0;
  GemmLaunchT<T> P{};
  P.nd = 0;
  InitTable<T> init{};
  long long init_rows = 0;
  const long long chunk = split_chunk<Cfg, T>(L, ksplit, maxparts);
  for (int i = 0; i < L.nd; ++i) {
    GemmDescT<T> d = L.d[i];
    if (d.batch <= 0 || d.M <= 0 || d.N <= 0) continue;
    if (d.beta != T(0) && d.Cin == nullptr) {
      set_error("gemm: beta != 0 requires Cin");
      return kBadArgument;
    }
    const bool inplace = d.Cin == d.C && d.beta == T(1) && d.ldcin == d.ldc &&
                         d.strideCin == d.strideC;
    if (d.atomic && !inplace) {
      set_error("gemm: an atomic descriptor must accumulate in place (Cin == C, beta == 1)");
      return kBadArgument;
    }
    const long long K = total_k(d);
    long long parts = chunk > 0 && !d.nosplit ? (K + chunk - 1) / chunk : 1;
    parts = parts < 1 ? 1 : (parts > maxparts ? maxparts : parts);
    const int ks = (int)parts;
    d.kchunk = (ceil_div((int)K, ks) + Cfg::BK - 1) / Cfg::BK * Cfg::BK;
    if (d.kchunk < Cfg::BK) d.kchunk = Cfg::BK;
    d.ksplit = ks > 1 ? ceil_div((int)K, d.kchunk) : 1;
    if (d.ksplit == 1) d.kchunk = (int)(K > 0 ? K : 1);
    if (d.ksplit > 1 && !inplace) {
      InitJob<T>& J = init.j[init.n++];
      J.C = d.C;
      J.Cin = d.Cin;
      J.sC = d.strideC;
      J.sCin = d.strideCin;
      J.ldc = d.ldc;
      J.ldcin = d.ldcin;
      J.M = d.M;
      J.N = d.N;
      J.beta = d.beta;
      J.row_begin = (int)init_rows;
      init_rows += (long long)d.batch * d.M;
      SSLAB_REQUIRE(init_rows <= 0x7fffffffLL, "gemm: split-K init table too large");
    }
    d.tiles_m = ceil_div(d.M, Cfg::BM);
    d.tiles_n = ceil_div(d.N, Cfg::BN);
    d.tile_begin = total;
    long long add = (long long)d.tiles_m * d.tiles_n * d.batch * d.ksplit;
    if (add + total > 0x7fffffffLL) {
      set_error("gemm: too many tiles in one launch");
      return kBadArgument;
    }
    total += (int)add;
    P.d[P.nd++] = d;
  }
  if (total == 0) return kOk;
  if (init.n > 0) {
    init.rows = (int)init_rows;
    const int grid = init.rows < 65536 ? init.rows : 65536;
    hipLaunchKernelGGL(k_scale_init<T>, dim3(grid), dim3(256), 0, stream, init);
    SSLAB_HIP_CHECK(hipGetLastError());
  }
  P.xcd_chunk = xcd_chunk(P, total);
  if constexpr (kADirect && ADGeom<Cfg, T>::ok) {
    int nmax = 0;
    for (int i = 0; i < P.nd; ++i) nmax = P.d[i].N > nmax ? P.d[i].N : nmax;
    if (nmax <= Cfg::BN && fast_kind<T>(P) == 2) {
      hipLaunchKernelGGL((k_gemm_adirect<Cfg, T>), dim3(total), dim3(Cfg::NT), 0, stream, P,
                         total);
      SSLAB_HIP_CHECK(hipGetLastError());
      return kOk;
    }
  }
  if constexpr (FastGeom<Cfg, T>::ok) {
    const int kind = fast_kind<T>(P);
    if (kind == 2) {
      static const bool pipe = [] {
        const char* e = getenv("SSLAB_GEMM_PIPE");
        return !(e && e[0] == '0');
      }();
      // the second register set spills for FP64 tiles of more than 128 x 64
      constexpr bool fits = !(sizeof(T) == 8 && Cfg::BM * Cfg::BN > 128 * 64);
      if (pipe && fits)
        hipLaunchKernelGGL((k_gemm_fast<Cfg, T, fits ? 2 : 1>), dim3(total), dim3(Cfg::NT), 0,
                           stream, P, total);
      else
        hipLaunchKernelGGL((k_gemm_fast<Cfg, T, 1>), dim3(total), dim3(Cfg::NT), 0, stream, P,
                           total);
      SSLAB_HIP_CHECK(hipGetLastError());
      return kOk;
    }
    if (kind == 1) {
      hipLaunchKernelGGL((k_gemm_fast<Cfg, T, 0>), dim3(total), dim3(Cfg::NT), 0, stream, P,
                         total);
      SSLAB_HIP_CHECK(hipGetLastError());
      return kOk;
    }
  }
  hipLaunchKernelGGL((k_gemm<Cfg, T>), dim3(total), dim3(Cfg::NT), 0, stream, P, total);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

template <typename T>
static int launch_id(int id, GemmLaunchT<T>& L, hipStream_t stream, int ks, int maxparts = 64) {
  switch (id) {
    case 0: return launch_cfg<Cfg128, T>(L, stream, ks, maxparts);
    case 1: return launch_cfg<Cfg128x64, T>(L, stream, ks, maxparts);
    case 2: return launch_cfg<Cfg64, T>(L, stream, ks, maxparts);
    case 3: return launch_cfg<Cfg64s, T>(L, stream, ks, maxparts);
    case 4: return launch_cfg<Cfg64x32, T>(L, stream, ks, maxparts);
    case 5: return launch_cfg<Cfg32, T>(L, stream, ks, maxparts);
    case 6: return launch_cfg<Cfg32x64, T>(L, stream, ks, maxparts);
    case 7: return launch_cfg<Cfg64x128, T>(L, stream, ks, maxparts);
    case 8: return launch_cfg<Cfg32x16, T>(L, stream, ks, maxparts);
    case 9: return launch_cfg<Cfg32w, T>(L, stream, ks, maxparts);
    case 10: return launch_cfg<CfgAD32, T, true>(L, stream, ks, maxparts);
    case 11: return launch_cfg<CfgAD64, T, true>(L, stream, ks, maxparts);
    case 12: return launch_cfg<Cfg32k, T>(L, stream, ks, maxparts);
    case 13: return launch_cfg<Cfg32k64, T>(L, stream, ks, maxparts);
    default: return launch_cfg<Cfg64x32k, T>(L, stream, ks, maxparts);
  }
}

static thread_local GemmPolicy g_policy;

int gemm_set_policy(const GemmPolicy& p) {
  SSLAB_REQUIRE(p.target_wgs >= 1 && p.split_wgs >= 1 && p.max_ksplit >= 1 && p.min_split_k >= 0,
                "gemm policy: counts must be positive");
  SSLAB_REQUIRE(p.cfg_mid >= 0 && p.cfg_mid < kNumCfg && p.cfg_split >= 0 &&
                    p.cfg_split < kNumCfg && p.max_tile >= 0 && p.max_tile <= 2,
                "gemm policy: tile id out of range");
  g_policy = p;
  return kOk;
}

GemmPolicy gemm_get_policy() { return g_policy; }

template <typename T>
int gemm_launch(GemmLaunchT<T>& L, hipStream_t stream) {
  SSLAB_REQUIRE(L.nd >= 0 && L.nd <= kMaxDesc, "gemm: bad descriptor count %d", L.nd);
  long long kmax = 0;
  for (int i = 0; i < L.nd; ++i) {
    const GemmDescT<T>& d = L.d[i];
    SSLAB_REQUIRE(d.nterms >= 1 && d.nterms <= kMaxTerms, "gemm: bad term count");
    for (int t = 0; t < d.nterms; ++t) {
      const GemmTermT<T>& Tm = d.term[t];
      SSLAB_REQUIRE(Tm.A && Tm.B, "gemm: null operand");
      SSLAB_REQUIRE(Tm.K >= 0, "gemm: negative K");
    }
    SSLAB_REQUIRE(total_k(d) < 0x7fffffffLL, "gemm: K too large");
    if (d.batch > 0 && total_k(d) > kmax) kmax = total_k(d);
  }
  // The largest tile that still gives the chip (256 CUs) policy.target_wgs workgroups.
  // Launches with few output tiles but long K (deep red-black levels, few right-hand
  // sides) split K over several workgroups instead, so every level fills the chip;
  // short-K launches fall back to smaller tiles.
  // Few right-hand sides (N <= 32) use the 16- and 32-column tiles, so the MFMA work
  // padded onto empty columns stays small and the launch streams A at HBM rate.
  const GemmPolicy& P = g_policy;
  const long long target = P.target_wgs;
  int nmax = 0;
  for (int i = 0; i < L.nd; ++i)
    if (L.d[i].batch > 0 && L.d[i].N > nmax) nmax = L.d[i].N;
  if (nmax > 64 && P.max_tile == 0 && count_tiles(L, 128, 128) >= target)
    return launch_id<T>(0, L, stream, 1);
  if (nmax > 32 && P.max_tile <= 1 && count_tiles(L, 128, 64) >= target)
    return launch_id<T>(1, L, stream, 1);
  const int mid = nmax <= 16 ? 8 : (nmax <= 32 ? 9 : P.cfg_mid);
  const int spl = nmax <= 16 ? 8 : (nmax <= 32 ? 9 : P.cfg_split);
  if (count_tiles(L, kCfgBM[mid], kCfgBN[mid]) >= target) return launch_id<T>(mid, L, stream, 1);
  if (kmax >= P.min_split_k) {
    // FP32 products split K into at most 4 parts: every part is rounded to FP32 before
    // the atomic add, and the FP32 red-black solve is the preconditioner of an FP64
    // refinement whose convergence follows its accuracy (C2 FP32 solve: 9e-5 with up to
    // 64 parts, 5.6e-5 with 4, 4e-5 unsplit; one refinement step then reaches 1.2e-8,
    // 4.4e-9 and 2e-9, and the 4-part cap keeps most of the split's speed).
    const long long tn = count_tiles(L, kCfgBM[spl], kCfgBN[spl]);
    const long long want = (P.split_wgs + tn - 1) / (tn > 0 ? tn : 1);
    const int cap = std::is_same<T, float>::value ? (P.max_ksplit < 4 ? P.max_ksplit : 4)
                                                  : P.max_ksplit;
    const int ks = (int)(want < cap ? want : cap);
    return launch_id<T>(spl, L, stream, ks, cap);
  }
  if (nmax <= 32) return launch_id<T>(mid, L, stream, 1);
  if (count_tiles(L, 64, 32) >= target) return launch_id<T>(4, L, stream, 1);
  return launch_id<T>(5, L, stream, 1);
}

// ---- dependency-ordered launch (k_gemm_chain) ------------------------------------------

template <class Cfg, typename T>
static int chain_cfg(GemmChainT<T>& L, hipStream_t stream, int ksplit, int maxparts) {
  if constexpr (!FastGeom<Cfg, T>::ok) {
    set_error("gemm chain: tile without the fast kernel");
    return kBadArgument;
  } else {
    const long long chunk = ksplit > 1 ? split_chunk<Cfg, T>(L, ksplit, maxparts) : 0;
    int total = 0;
    for (int i = 0; i < kChainBlocks; ++i) L.need[i] = 0;
    for (int i = 0; i < L.nd; ++i) {
      GemmDescT<T>& d = L.d[i];
      SSLAB_REQUIRE(d.batch > 0 && d.M > 0 && d.N > 0, "gemm chain: empty descriptor");
      const bool inplace = d.Cin == d.C && d.beta == T(1) && d.ldcin == d.ldc &&
                           d.strideCin == d.strideC;
      // only in-place products split K (their parts add atomically into C); any other is
      // written once by plain stores, C = beta Cin + sum of terms, with no init pass
      const bool single = d.nosplit || !inplace;
      const long long K = total_k(d);
      long long parts = chunk > 0 && !single ? (K + chunk - 1) / chunk : 1;
      parts = parts < 1 ? 1 : (parts > maxparts ? maxparts : parts);
      d.kchunk = (ceil_div((int)K, (int)parts) + Cfg::BK - 1) / Cfg::BK * Cfg::BK;
      if (d.kchunk < Cfg::BK) d.kchunk = Cfg::BK;
      d.ksplit = parts > 1 ? ceil_div((int)K, d.kchunk) : 1;
      if (d.ksplit == 1) d.kchunk = (int)(K > 0 ? K : 1);
      SSLAB_REQUIRE(d.ksplit == 1 || inplace, "gemm chain: split-K needs an in-place product");
      d.tiles_m = ceil_div(d.M, Cfg::BM);
      d.tiles_n = ceil_div(d.N, Cfg::BN);
      d.tile_begin = total;
      const int per = d.tiles_m * d.tiles_n * d.ksplit;
      total += per * d.batch;
      for (int t = 0; t < d.nterms; ++t)
        SSLAB_REQUIRE(L.dep1[i][t] == 0 ||
                          (L.dep1[i][t] >= 1 &&
                           L.dep1[i][t] - 1 + (long long)(d.batch - 1) * L.dep_stride[i][t] <
                               kChainBlocks),
                      "gemm chain: dependency block out of range");
      if (L.sig1[i] != 0) {
        SSLAB_REQUIRE(L.sig1[i] >= 1 &&
                          L.sig1[i] - 1 + (long long)(d.batch - 1) * L.sig_stride[i] < kChainBlocks,
                      "gemm chain: signalled block out of range");
        for (int b = 0; b < d.batch; ++b) L.need[L.sig1[i] - 1 + b * L.sig_stride[i]] += per;
      }
    }
    // a tile may only wait on blocks written by earlier descriptors of the launch
    for (int i = 0; i < L.nd; ++i)
      for (int t = 0; t < L.d[i].nterms; ++t) {
        if (L.dep1[i][t] == 0) continue;
        for (int j = i; j < L.nd; ++j)
          for (int b = 0; b < L.d[i].batch && L.sig1[j] != 0; ++b)
            for (int c = 0; c < L.d[j].batch; ++c)
              SSLAB_REQUIRE(L.dep1[i][t] - 1 + b * L.dep_stride[i][t] !=
                                L.sig1[j] - 1 + c * L.sig_stride[j],
                            "gemm chain: descriptor %d waits on a block written by %d", i, j);
      }
    if (total == 0) return kOk;
    const int kind = fast_kind<T>(L);
    if (kind == 2)
      hipLaunchKernelGGL((k_gemm_chain<Cfg, T, 2>), dim3(total), dim3(Cfg::NT), 0, stream, L,
                         total);
    else
      hipLaunchKernelGGL((k_gemm_chain<Cfg, T, 0>), dim3(total), dim3(Cfg::NT), 0, stream, L,
                         total);
    SSLAB_HIP_CHECK(hipGetLastError());
    return kOk;
  }
}

template <typename T>
int gemm_chain_launch(GemmChainT<T>& L, hipStream_t stream) {
  SSLAB_REQUIRE(L.nd >= 1 && L.nd <= kChainDesc, "gemm chain: bad descriptor count %d", L.nd);
  SSLAB_REQUIRE(L.cnt && L.ctl && L.spin_limit > 0, "gemm chain: missing counters");
  for (int i = 0; i < L.nd; ++i) {
    const GemmDescT<T>& d = L.d[i];
    SSLAB_REQUIRE(d.nterms >= 1 && d.nterms <= kMaxTerms, "gemm chain: bad term count");
    for (int t = 0; t < d.nterms; ++t)
      SSLAB_REQUIRE(d.term[t].A && d.term[t].B && d.term[t].K >= 0, "gemm chain: bad term");
  }
  const GemmPolicy& P = g_policy;
  const int ks = P.max_ksplit < 1 ? 1 : P.max_ksplit;
  long long tiles = 0;
  for (int i = 0; i < L.nd; ++i)
    tiles += (long long)ceil_div(L.d[i].M, kCfgBM[P.cfg_split]) *
             ceil_div(L.d[i].N, kCfgBN[P.cfg_split]) * L.d[i].batch;
  const long long want = (P.split_wgs + tiles - 1) / (tiles > 0 ? tiles : 1);
  const int split = (int)(want < ks ? want : ks);
  switch (P.cfg_split) {
    case 2: return chain_cfg<Cfg64, T>(L, stream, split, ks);
    case 3: return chain_cfg<Cfg64s, T>(L, stream, split, ks);
    case 4: return chain_cfg<Cfg64x32, T>(L, stream, split, ks);
    case 5: return chain_cfg<Cfg32, T>(L, stream, split, ks);
    case 9: return chain_cfg<Cfg32w, T>(L, stream, split, ks);
    case 12: return chain_cfg<Cfg32k, T>(L, stream, split, ks);
    case 13: return chain_cfg<Cfg32k64, T>(L, stream, split, ks);
    case 14: return chain_cfg<Cfg64x32k, T>(L, stream, split, ks);
    default:
      set_error("gemm chain: tile id %d is not available for dependency-ordered launches",
                P.cfg_split);
      return kBadArgument;
  }
}

template int gemm_chain_launch<double>(GemmChainT<double>&, hipStream_t);

template <typename T>
int gemm_strided(int transA, int transB, int M, int N, int K, typename same_t<T>::type alpha,
                 const T* A, int lda, long long sA, const T* B, int ldb, long long sB,
                 typename same_t<T>::type beta, T* C, int ldc, long long sC, int batch,
                 hipStream_t stream) {
  GemmLaunchT<T> L{};
  L.nd = 1;
  L.d[0] = gemm_desc<T>(M, N, batch, C, ldc, sC, beta != T(0) ? C : nullptr, ldc, sC, beta);
  gemm_add_term<T>(L.d[0], gemm_term<T>(A, lda, sA, transA, B, ldb, sB, transB, K, alpha));
  return gemm_launch<T>(L, stream);
}

template int gemm_launch<double>(GemmLaunchT<double>&, hipStream_t);
template int gemm_launch<float>(GemmLaunchT<float>&, hipStream_t);
template int gemm_strided<double>(int, int, int, int, int, double, const double*, int, long long,
                                  const double*, int, long long, double, double*, int, long long,
                                  int, hipStream_t);
template int gemm_strided<float>(int, int, int, int, int, float, const float*, int, long long,
                                 const float*, int, long long, float, float*, int, long long, int,
                                 hipStream_t);

// ---- element-wise helpers -------------------------------------------------------------

namespace {

template <typename Tin, typename Tout>
__global__ void k_cast(long long n, const Tin* __restrict__ x, Tout* __restrict__ y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = (Tout)x[i];
}

__global__ void k_acc(long long n, const float* __restrict__ x, double* __restrict__ y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] += (double)x[i];
}

__global__ void k_sub(long long n, const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ z) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    z[i] = a[i] - b[i];
}

int grid_for(long long n) {
  long long g = (n + 255) / 256;
  return (int)(g < 8192 ? (g > 0 ? g : 1) : 8192);
}

template <typename T>
__global__ void k_zero_2d(long long rows, int cols, T* __restrict__ X, int ld) {
  const long long n = rows * cols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const long long i = e / cols;
    X[i * ld + (e - i * cols)] = T(0);
  }
}

}  // namespace

template <typename T>
int zero_2d(long long rows, int cols, T* X, int ld, hipStream_t st) {
  if (rows <= 0 || cols <= 0) return kOk;
  SSLAB_REQUIRE(ld >= cols && X != nullptr, "zero_2d: bad buffer");
  hipLaunchKernelGGL(k_zero_2d<T>, dim3(grid_for(rows * cols)), dim3(256), 0, st, rows, cols, X, ld);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

template int zero_2d<double>(long long, int, double*, int, hipStream_t);
template int zero_2d<float>(long long, int, float*, int, hipStream_t);

int cast_f64_to_f32(long long n, const double* x, float* y, hipStream_t st) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL((k_cast<double, float>), dim3(grid_for(n)), dim3(256), 0, st, n, x, y);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

int cast_f32_to_f64(long long n, const float* x, double* y, hipStream_t st) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL((k_cast<float, double>), dim3(grid_for(n)), dim3(256), 0, st, n, x, y);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

int acc_f32_into_f64(long long n, const float* x, double* y, hipStream_t st) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(k_acc, dim3(grid_for(n)), dim3(256), 0, st, n, x, y);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

int sub_f64(long long n, const double* a, const double* b, double* z, hipStream_t st) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(k_sub, dim3(grid_for(n)), dim3(256), 0, st, n, a, b, z);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

}  // namespace sslab
