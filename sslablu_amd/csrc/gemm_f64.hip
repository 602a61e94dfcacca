// Grouped / batched FP64 GEMM for gfx950 (MFMA v_mfma_f64_16x16x4_f64).
//
// Tiling: a workgroup computes a BM x BN tile of C with WM x WN wavefronts (64
// lanes each); a wavefront owns a (BM/WM) x (BN/WN) sub-tile made of 16x16 MFMA
// fragments.  K is consumed in BK-wide slabs staged through LDS, k-major with a
// pad so that the 16-lane row groups of one MFMA operand read fall in disjoint
// bank ranges.  The next slab is prefetched into registers while the current
// one is multiplied.
//
// Operand lane layout of v_mfma_f64_16x16x4_f64 (one double per lane):
//   A (16x4):  lane l holds A[l % 16][l / 16]
//   B (4x16):  lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, item v holds D[(l/16) + 4*v][l % 16]
// (measured on gfx950; unlike the FP32 16x16 MFMAs, whose item v is row 4*(l/16)+v).
#include "gemm_f64.hpp"

namespace sslab {

typedef double d4_t __attribute__((ext_vector_type(4)));

template <int BM_, int BN_, int BK_, int WM_, int WN_>
struct GemmCfg {
  static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
  static constexpr int NT = WM * WN * 64;
  static constexpr int WTM = BM / WM, WTN = BN / WN;  // wavefront tile
  static constexpr int FM = WTM / 16, FN = WTN / 16;   // MFMA fragments per wave
  // k-major LDS rows padded so consecutive k rows start 128 bytes apart mod 256.
  static constexpr int LDA_S = BM + 16;
  static constexpr int LDB_S = BN + 16;
  static constexpr int A_PER_T = BM * BK / NT;
  static constexpr int B_PER_T = BK * BN / NT;
  static_assert(BM % (WM * 16) == 0 && BN % (WN * 16) == 0, "tile/wave mismatch");
  static_assert(BK % 4 == 0, "BK must be a multiple of the MFMA K (4)");
  static_assert((BM * BK) % NT == 0 && (BK * BN) % NT == 0, "load split");
};

// Map the hardware block id so that contiguous ranges of logical tiles land on the
// same XCD (the dispatcher assigns workgroups round-robin to the 8 XCDs, each with
// its own L2).  Bijective on [0, n).
__device__ __forceinline__ int xcd_remap(int bid, int n) {
  const int q = n >> 3;
  if (bid < (q << 3)) return (bid & 7) * q + (bid >> 3);
  return bid;
}

template <class Cfg>
__device__ __forceinline__ void load_a_regs(const GemmTerm& t, const double* A, int M,
                                            int row0, int k0, int tid,
                                            double (&ra)[Cfg::A_PER_T]) {
  const int K = t.K;
  if (!t.transA) {
    // contiguous along k
#pragma unroll
    for (int r = 0; r < Cfg::A_PER_T; ++r) {
      const int e = tid + r * Cfg::NT;
      const int k = e % Cfg::BK, i = e / Cfg::BK;
      const int gi = row0 + i, gk = k0 + k;
      ra[r] = (gi < M && gk < K) ? A[(long long)gi * t.lda + gk] : 0.0;
    }
  } else {
#pragma unroll
    for (int r = 0; r < Cfg::A_PER_T; ++r) {
      const int e = tid + r * Cfg::NT;
      const int i = e % Cfg::BM, k = e / Cfg::BM;
      const int gi = row0 + i, gk = k0 + k;
      ra[r] = (gi < M && gk < K) ? A[(long long)gk * t.lda + gi] : 0.0;
    }
  }
}

template <class Cfg>
__device__ __forceinline__ void store_a_lds(const GemmTerm& t, int tid, double alpha,
                                            const double (&ra)[Cfg::A_PER_T], double* As) {
  if (!t.transA) {
#pragma unroll
    for (int r = 0; r < Cfg::A_PER_T; ++r) {
      const int e = tid + r * Cfg::NT;
      const int k = e % Cfg::BK, i = e / Cfg::BK;
      As[k * Cfg::LDA_S + i] = alpha * ra[r];
    }
  } else {
#pragma unroll
    for (int r = 0; r < Cfg::A_PER_T; ++r) {
      const int e = tid + r * Cfg::NT;
      const int i = e % Cfg::BM, k = e / Cfg::BM;
      As[k * Cfg::LDA_S + i] = alpha * ra[r];
    }
  }
}

template <class Cfg>
__device__ __forceinline__ void load_b_regs(const GemmTerm& t, const double* B, int N,
                                            int col0, int k0, int tid,
                                            double (&rb)[Cfg::B_PER_T]) {
  const int K = t.K;
  if (!t.transB) {
    // contiguous along j
#pragma unroll
    for (int r = 0; r < Cfg::B_PER_T; ++r) {
      const int e = tid + r * Cfg::NT;
      const int j = e % Cfg::BN, k = e / Cfg::BN;
      const int gj = col0 + j, gk = k0 + k;
      rb[r] = (gj < N && gk < K) ? B[(long long)gk * t.ldb + gj] : 0.0;
    }
  } else {
#pragma unroll
    for (int r = 0; r < Cfg::B_PER_T; ++r) {
      const int e = tid + r * Cfg::NT;
      const int k = e % Cfg::BK, j = e / Cfg::BK;
      const int gj = col0 + j, gk = k0 + k;
      rb[r] = (gj < N && gk < K) ? B[(long long)gj * t.ldb + gk] : 0.0;
    }
  }
}

template <class Cfg>
__device__ __forceinline__ void store_b_lds(const GemmTerm& t, int tid,
                                            const double (&rb)[Cfg::B_PER_T], double* Bs) {
  if (!t.transB) {
#pragma unroll
    for (int r = 0; r < Cfg::B_PER_T; ++r) {
      const int e = tid + r * Cfg::NT;
      const int j = e % Cfg::BN, k = e / Cfg::BN;
      Bs[k * Cfg::LDB_S + j] = rb[r];
    }
  } else {
#pragma unroll
    for (int r = 0; r < Cfg::B_PER_T; ++r) {
      const int e = tid + r * Cfg::NT;
      const int k = e % Cfg::BK, j = e / Cfg::BK;
      Bs[k * Cfg::LDB_S + j] = rb[r];
    }
  }
}

template <class Cfg>
__global__ void __launch_bounds__(Cfg::NT) k_gemm_f64(const GemmLaunch L, int total_tiles) {
  __shared__ double As[Cfg::BK * Cfg::LDA_S];
  __shared__ double Bs[Cfg::BK * Cfg::LDB_S];

  const int tile = xcd_remap(blockIdx.x, total_tiles);

  // Select the descriptor with a fully unrolled scan (constant kernarg offsets).
  GemmDesc D = L.d[0];
#pragma unroll
  for (int d = 1; d < kMaxDesc; ++d)
    if (d < L.nd && tile >= L.d[d].tile_begin) D = L.d[d];

  const int per_batch = D.tiles_m * D.tiles_n;
  int local = tile - D.tile_begin;
  const int b = local / per_batch;
  local -= b * per_batch;
  const int tm = local / D.tiles_n;
  const int tn = local - tm * D.tiles_n;
  const int row0 = tm * Cfg::BM, col0 = tn * Cfg::BN;
  const int M = D.M, N = D.N;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / Cfg::WN, wn = wave % Cfg::WN;
  const int wrow = wm * Cfg::WTM, wcol = wn * Cfg::WTN;

  d4_t acc[Cfg::FM][Cfg::FN];
#pragma unroll
  for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
    for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

  double ra[Cfg::A_PER_T], rb[Cfg::B_PER_T];

#pragma unroll
  for (int ti = 0; ti < kMaxTerms; ++ti) {
    if (ti >= D.nterms) break;
    const GemmTerm T = D.term[ti];
    const double* A = T.A + (long long)b * T.strideA;
    const double* B = T.B + (long long)b * T.strideB;
    const int K = T.K;
    if (K <= 0) continue;
    load_a_regs<Cfg>(T, A, M, row0, 0, tid, ra);
    load_b_regs<Cfg>(T, B, N, col0, 0, tid, rb);
    for (int k0 = 0; k0 < K; k0 += Cfg::BK) {
      __syncthreads();
      store_a_lds<Cfg>(T, tid, T.alpha, ra, As);
      store_b_lds<Cfg>(T, tid, rb, Bs);
      __syncthreads();
      if (k0 + Cfg::BK < K) {
        load_a_regs<Cfg>(T, A, M, row0, k0 + Cfg::BK, tid, ra);
        load_b_regs<Cfg>(T, B, N, col0, k0 + Cfg::BK, tid, rb);
      }
#pragma unroll
      for (int kk = 0; kk < Cfg::BK; kk += 4) {
        const int kr = kk + (lane >> 4);
        double af[Cfg::FM], bf[Cfg::FN];
#pragma unroll
        for (int i = 0; i < Cfg::FM; ++i)
          af[i] = As[kr * Cfg::LDA_S + wrow + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j)
          bf[j] = Bs[kr * Cfg::LDB_S + wcol + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
          for (int j = 0; j < Cfg::FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
      }
    }
  }

  double* C = D.C + (long long)b * D.strideC;
  const double* Cin = D.Cin ? D.Cin + (long long)b * D.strideCin : nullptr;
  const double beta = D.beta;
  const bool use_cin = (Cin != nullptr) && (beta != 0.0);
#pragma unroll
  for (int i = 0; i < Cfg::FM; ++i) {
#pragma unroll
    for (int j = 0; j < Cfg::FN; ++j) {
      const int col = col0 + wcol + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = row0 + wrow + i * 16 + (lane >> 4) + 4 * v;
        if (row < M && col < N) {
          double val = acc[i][j][v];
          if (use_cin) val += beta * Cin[(long long)row * D.ldcin + col];
          C[(long long)row * D.ldc + col] = val;
        }
      }
    }
  }
}

using Cfg64 = GemmCfg<64, 64, 16, 2, 2>;
using Cfg128 = GemmCfg<128, 128, 16, 2, 2>;
using Cfg128x64 = GemmCfg<128, 64, 16, 2, 2>;

template <class Cfg>
static int launch_cfg(GemmLaunch& L, hipStream_t stream) {
  int total = 0;
  GemmLaunch P{};
  P.nd = 0;
  for (int i = 0; i < L.nd; ++i) {
    GemmDesc d = L.d[i];
    if (d.batch <= 0 || d.M <= 0 || d.N <= 0) continue;
    if (d.beta != 0.0 && d.Cin == nullptr) {
      set_error("gemm: beta != 0 requires Cin");
      return kBadArgument;
    }
    d.tiles_m = ceil_div(d.M, Cfg::BM);
    d.tiles_n = ceil_div(d.N, Cfg::BN);
    d.tile_begin = total;
    long long add = (long long)d.tiles_m * d.tiles_n * d.batch;
    if (add + total > 0x7fffffffLL) {
      set_error("gemm: too many tiles in one launch");
      return kBadArgument;
    }
    total += (int)add;
    P.d[P.nd++] = d;
  }
  if (total == 0) return kOk;
  hipLaunchKernelGGL(k_gemm_f64<Cfg>, dim3(total), dim3(Cfg::NT), 0, stream, P, total);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

int gemm_launch(GemmLaunch& L, hipStream_t stream) {
  SSLAB_REQUIRE(L.nd >= 0 && L.nd <= kMaxDesc, "gemm: bad descriptor count %d", L.nd);
  // Pick the tile shape from the largest descriptor: big square products use
  // 128x128, tall-skinny (few right-hand sides) 128x64, small ones 64x64.
  long long work = 0;
  int maxM = 0, maxN = 0;
  for (int i = 0; i < L.nd; ++i) {
    const GemmDesc& d = L.d[i];
    SSLAB_REQUIRE(d.nterms >= 1 && d.nterms <= kMaxTerms, "gemm: bad term count");
    for (int t = 0; t < d.nterms; ++t) {
      const GemmTerm& T = d.term[t];
      SSLAB_REQUIRE(T.A && T.B, "gemm: null operand");
      SSLAB_REQUIRE(T.K >= 0, "gemm: negative K");
    }
    if (d.batch > 0) {
      work += (long long)d.M * d.N * d.batch;
      if (d.M > maxM) maxM = d.M;
      if (d.N > maxN) maxN = d.N;
    }
  }
  const long long tiles128 = work / (128 * 128);
  if (maxM >= 128 && maxN >= 128 && tiles128 >= 256) return launch_cfg<Cfg128>(L, stream);
  if (maxM >= 128 && maxN <= 96 && work / (128 * 64) >= 256) return launch_cfg<Cfg128x64>(L, stream);
  return launch_cfg<Cfg64>(L, stream);
}

int gemm_strided(int transA, int transB, int M, int N, int K, double alpha, const double* A,
                 int lda, long long sA, const double* B, int ldb, long long sB, double beta,
                 double* C, int ldc, long long sC, int batch, hipStream_t stream) {
  GemmLaunch L{};
  L.nd = 1;
  L.d[0] = gemm_desc(M, N, batch, C, ldc, sC, beta != 0.0 ? C : nullptr, ldc, sC, beta);
  gemm_add_term(L.d[0], gemm_term(A, lda, sA, transA, B, ldb, sB, transB, K, alpha));
  return gemm_launch(L, stream);
}

}  // namespace sslab
