// MFMA GEMMs for gfx950 (CDNA4), row-major, strided-batched, with operand
// transposes folded into the global->LDS staging (no separate permute passes).
//
// FP64 / FP32: v_mfma_{f64,f32}_16x16x4 -- 256 threads = 4 waves (2x2), each wave
// owns a (BM/2)x(BN/2) sub-tile built from 16x16 MFMA blocks; K is staged in
// LDS slabs of 16 with register prefetch of the next slab (double-buffered LDS,
// one barrier per slab).  Lane layouts of the 16x16x4 instructions:
//   A operand: lane l holds A[l%16][l/16]    B operand: lane l holds B[l/16][l%16]
//   D (f32):   lane l, item r holds D[4*(l/16)+r][l%16]
//   D (f64):   lane l, item r holds D[l/16+4*r][l%16]
// BF16: v_mfma_f32_16x16x32_bf16 -- A: lane l holds A[l%16][8*(l/16)+j], j<8;
// B: B[8*(l/16)+j][l%16]; D as above.  LDS holds K-contiguous rows so every
// fragment is one 16-byte LDS read.
#include "internal.h"

namespace nsmi {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

template <typename T> struct Acc;
template <> struct Acc<double> { typedef d4 type; };
template <> struct Acc<float> { typedef f4 type; };

__device__ inline d4 mfma16(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline f4 mfma16(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row of item r of lane group kl = lane/16 in a 16x16 MFMA result block
// (measured on gfx950): f64 16x16x4 interleaves rows kl + 4r; f32 16x16x4
// and bf16 16x16x32 hold four consecutive rows 4*kl + r.
template <typename T> __device__ inline int out_row(int kl, int r);
template <> __device__ inline int out_row<double>(int kl, int r) { return kl + 4 * r; }
template <> __device__ inline int out_row<float>(int kl, int r) { return 4 * kl + r; }

template <typename T>
struct GemmArgs {
  const T* A;
  const T* B;
  T* C;
  int64_t M, N, K, lda, ldb, ldc, sA, sB, sC;
  T alpha, beta;
  int tiles_m, tiles_n;
};

// XCD-aware tile order: consecutive workgroup ids are dealt round-robin to the
// 8 XCDs, so remap them such that each XCD works on a contiguous band of tiles
// (neighbouring tiles share A rows / B columns in that XCD's L2).
__device__ inline int xcd_remap(int wg, int nwg) {
  constexpr int kXcd = 8;
  if (nwg < 2 * kXcd || (nwg % kXcd) != 0) return wg;
  return (wg % kXcd) * (nwg / kXcd) + wg / kXcd;
}

template <typename T, int BM, int BN, bool TA, bool TB, int ROLE>
__global__ __launch_bounds__(256) void k_gemm_mfma(GemmArgs<T> g) {
  constexpr int BK = 16, PAD = 4;
  constexpr int WTM = BM / 2, WTN = BN / 2;
  constexpr int FM = WTM / 16, FN = WTN / 16;
  constexpr int EA = BM * BK / 256, EB = BN * BK / 256;
  typedef typename Acc<T>::type AccT;
  __shared__ T sA[2][BK][BM + PAD];
  __shared__ T sB[2][BK][BN + PAD];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int nwg = g.tiles_m * g.tiles_n;
  const int tile = xcd_remap(blockIdx.x, nwg);
  const int tm = tile / g.tiles_n, tn = tile % g.tiles_n;
  const int64_t m0 = int64_t(tm) * BM, n0 = int64_t(tn) * BN;
  const int64_t bz = blockIdx.z;
  const T* __restrict__ A = g.A + bz * g.sA;
  const T* __restrict__ B = g.B + bz * g.sB;
  T* __restrict__ C = g.C + bz * g.sC;

  T ra[EA], rb[EB];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int j = 0; j < EA; ++j) {
      const int e = tid * EA + j;
      const int mm = TA ? (e % BM) : (e / BK);
      const int kk = TA ? (e / BM) : (e % BK);
      const int64_t gm = m0 + mm, gk = k0 + kk;
      T v = T(0);
      if (gm < g.M && gk < g.K) v = TA ? A[gk * g.lda + gm] : A[gm * g.lda + gk];
      ra[j] = v;
    }
#pragma unroll
    for (int j = 0; j < EB; ++j) {
      const int e = tid * EB + j;
      const int nn = TB ? (e / BK) : (e % BN);
      const int kk = TB ? (e % BK) : (e / BN);
      const int64_t gn = n0 + nn, gk = k0 + kk;
      T v = T(0);
      if (gn < g.N && gk < g.K) v = TB ? B[gn * g.ldb + gk] : B[gk * g.ldb + gn];
      rb[j] = v;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < EA; ++j) {
      const int e = tid * EA + j;
      const int mm = TA ? (e % BM) : (e / BK);
      const int kk = TA ? (e / BM) : (e % BK);
      sA[buf][kk][mm] = ra[j];
    }
#pragma unroll
    for (int j = 0; j < EB; ++j) {
      const int e = tid * EB + j;
      const int nn = TB ? (e / BK) : (e % BN);
      const int kk = TB ? (e % BK) : (e / BN);
      sB[buf][kk][nn] = rb[j];
    }
  };

  AccT acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = AccT{0, 0, 0, 0};

  const int64_t nk = (g.K + BK - 1) / BK;
  load(0);
  store(0);
  __syncthreads();
  const int kl = lane >> 4, ml = lane & 15;
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int buf = int(kt & 1);
    if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      T af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = sA[buf][ks * 4 + kl][wm * WTM + i * 16 + ml];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = sB[buf][ks * 4 + kl][wn * WTN + j * 16 + ml];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = mfma16(af[i], bf[j], acc[i][j]);
    }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gm = m0 + wm * WTM + i * 16 + out_row<T>(kl, r);
        const int64_t gn = n0 + wn * WTN + j * 16 + ml;
        if (gm < g.M && gn < g.N) {
          T v = g.alpha * acc[i][j][r];
          if (g.beta != T(0)) v += g.beta * C[gm * g.ldc + gn];
          C[gm * g.ldc + gn] = v;
        }
      }
}

template <typename T, int BM, int BN, int ROLE>
void launch_tile(nsmi_ctx* ctx, const GemmDesc& d, GemmArgs<T>& a) {
  a.tiles_m = int((d.M + BM - 1) / BM);
  a.tiles_n = int((d.N + BN - 1) / BN);
  dim3 grid(unsigned(a.tiles_m) * unsigned(a.tiles_n), 1, unsigned(d.batch));
  dim3 block(256);
  if (!d.transA && !d.transB)
    hipLaunchKernelGGL((k_gemm_mfma<T, BM, BN, false, false, ROLE>), grid, block, 0, ctx->stream, a);
  else if (!d.transA && d.transB)
    hipLaunchKernelGGL((k_gemm_mfma<T, BM, BN, false, true, ROLE>), grid, block, 0, ctx->stream, a);
  else if (d.transA && !d.transB)
    hipLaunchKernelGGL((k_gemm_mfma<T, BM, BN, true, false, ROLE>), grid, block, 0, ctx->stream, a);
  else
    hipLaunchKernelGGL((k_gemm_mfma<T, BM, BN, true, true, ROLE>), grid, block, 0, ctx->stream, a);
  NSMI_LAUNCH_CHECK();
}

template <typename T, int ROLE>
void launch_role(nsmi_ctx* ctx, const GemmDesc& d, GemmArgs<T>& a) {
  // 128x128 tiles when they fill the chip (>= 2 per CU), else 64x64
  const int64_t big = ((d.M + 127) / 128) * ((d.N + 127) / 128) * d.batch;
  if (big >= 512)
    launch_tile<T, 128, 128, ROLE>(ctx, d, a);
  else
    launch_tile<T, 64, 64, ROLE>(ctx, d, a);
}

template <typename T>
void gemm_generic(nsmi_ctx* ctx, const GemmDesc& d, T alpha, const T* A, const T* B, T beta, T* C, int role) {
  require(d.M >= 0 && d.N >= 0 && d.K >= 0 && d.batch >= 1, "gemm: bad sizes");
  require(d.tiles_ok(), "gemm: too many tiles");
  if (d.M == 0 || d.N == 0) return;
  GemmArgs<T> a{A, B, C, d.M, d.N, d.K, d.lda, d.ldb, d.ldc, d.sA, d.sB, d.sC, alpha, beta, 0, 0};
  switch (role) {
    case kRoleHeff1: launch_role<T, kRoleHeff1>(ctx, d, a); break;
    case kRoleHeff2: launch_role<T, kRoleHeff2>(ctx, d, a); break;
    case kRoleEnv: launch_role<T, kRoleEnv>(ctx, d, a); break;
    case kRoleTheta: launch_role<T, kRoleTheta>(ctx, d, a); break;
    default: launch_role<T, kRoleFact>(ctx, d, a); break;
  }
}

}  // namespace

bool GemmDesc::tiles_ok() const {
  const int64_t t = ((M + 63) / 64) * ((N + 63) / 64);
  return t < (int64_t(1) << 30) && batch < 65536;
}

void gemm_f64(nsmi_ctx* ctx, const GemmDesc& g, double alpha, const double* A, const double* B, double beta,
              double* C, int role) {
  gemm_generic<double>(ctx, g, alpha, A, B, beta, C, role);
}

void gemm_f32(nsmi_ctx* ctx, const GemmDesc& g, const float* A, const float* B, float* C, int role) {
  gemm_generic<float>(ctx, g, 1.0f, A, B, 0.0f, C, role);
}

}  // namespace nsmi
