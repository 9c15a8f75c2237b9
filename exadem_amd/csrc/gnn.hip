// gnn.hip — learned surrogate of K DEM steps on the contact neighbour list.
//
// Network (no counterpart in the reference; the CPU restatement is the test oracle's gnn module):
//   h0 = relu(x W_in^T + b_in)
//   layer l: [P|Q] = h W_AB^T;  z_ij = P_i + Q_j + e_ij W_E^T + b_E;  m_i = 1/8 sum_j relu(z_ij)
//   (outputs: mean contact force / (m0 a0) and torque / (m0 a0 r0) over the next K DEM steps)
//            U = [h|m] W_U^T + b_U;  h <- h + relu(U)
//   out = h W_out^T + b_out
// Dense products over particles (P|Q, U and their backward and weight-gradient products) run on
// MFMA (v_mfma_f32_32x32x16_bf16); the per-edge work (messages and their backward) is a segment sum
// over each particle's list entries, one wavefront per particle, two channels per lane. The backward
// of the Q side reads the reverse entry of every edge (its mirror slot), so no atomics are used and
// every reduction has a fixed order.
//
// HBM layout per network: node inputs x [n][8] bf16, edge inputs e [slots][16] bf16 (same slot index
// as the neighbour list), hidden h [n][128] bf16 (L+1 copies when training), P|Q [n][256] bf16,
// messages m [n][128] bf16, update pre-activations U [n][128] bf16, outputs [n][6] fp32.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "xdem_internal.h"

namespace xdem {
namespace gnn {

constexpr int H = XDEM_GNN_HIDDEN, L = XDEM_GNN_LAYERS, FN = 8, FE = 16, FO = 6, H2 = 2 * H;
constexpr float kMsgScale = 0.125f;
constexpr int kRedBlocks = 512;  // blocks of the grid-stride reduction kernels (partial sums)
// largest partial-sum need: 256 weight-gradient blocks of 128 x 256 floats
constexpr int64_t kPartialFloats = (int64_t)256 * 128 * 256 + 65536;

struct Off {
  int64_t W_in, b_in, W_AB[L], W_E[L], b_E[L], W_U[L], b_U[L], W_out, b_out, total;
};
inline Off offsets() {
  Off o{};
  int64_t p = 0;
  o.W_in = p; p += (int64_t)H * FN;
  o.b_in = p; p += H;
  for (int l = 0; l < L; l++) {
    o.W_AB[l] = p; p += (int64_t)H2 * H;
    o.W_E[l] = p; p += (int64_t)H * FE;
    o.b_E[l] = p; p += H;
    o.W_U[l] = p; p += (int64_t)H * H2;
    o.b_U[l] = p; p += H;
  }
  o.W_out = p; p += (int64_t)FO * H;
  o.b_out = p; p += FO;
  o.total = p;
  return o;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline float bf2f(uint16_t v) {
  uint32_t u = (uint32_t)v << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
// round to nearest even; NaN stays NaN
// float -> bfloat16, round to nearest even, on the hardware conversion (v_cvt_pk_bf16_f32)
typedef float float2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
__device__ inline uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
// the same rounding in integer arithmetic: cheaper than the hardware conversion inside kernels that
// keep the FP64 divide unit busy (k_features: 198 us against 265 us at 1M particles)
__device__ inline uint16_t f2bf_int(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline uint32_t pack_bf16_int(float lo, float hi) {
  return (uint32_t)f2bf_int(lo) | ((uint32_t)f2bf_int(hi) << 16);
}
// two values in one 32-bit word, `lo` in the low half
__device__ inline uint32_t pack_bf16(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(float2v{lo, hi}, bf16x2v));
}
__device__ inline float rbf(float f) { return bf2f(f2bf(f)); }

// ------------------------------------------------------------------------------------- inputs
struct FeatArgs {
  int64_t n;
  PState s;
  Grid g;
  NbrList nl;
  float r0, v0, ov;  // scales (inverse where noted in the kernel)
  uint16_t* x;       // [n][8]
  uint16_t* e;       // [slots][8]
};

// node inputs: ((R - r0)/r0 * 4, w R / v0, (R/r0)^3, 0, 0, 0); edge inputs for every list entry i -> j,
// with d = (r_j + image) - r_i, n = d/|d|, Rs = R_i + R_j, o = (Rs - |d|)/Rs * ov (overlap, > 0 in
// contact), c = [o > 0], v = relative surface velocity at the contact point
// [(v_j - R_j w_j x n) - (v_i + R_i w_i x n)] / v0, vn = v . n:
//   e0..2 = n, e3 = o, e4..6 = v, e7 = (R_j - R_i)/Rs * 4,
//   e8..10 = max(o, 0) n, e11..13 = c vn n, e14 = max(o, 0), e15 = c
__global__ __launch_bounds__(256) void k_features(FeatArgs a) {
  // XCD-aware block order (consecutive particle ranges share an XCD's L2 for the partner gathers)
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const Core ci = a.s.core[i];
  const V3 ri{ci.xr.x, ci.xr.y, ci.xr.z}, vi{ci.vm.x, ci.vm.y, ci.vm.z}, wi{ci.w.x, ci.w.y, ci.w.z};
  const double Ri = ci.xr.w;
  const float inv_v0 = 1.0f / a.v0;
  {
    uint16_t xv[8];
    const float rr = (float)(Ri / a.r0);
    xv[0] = f2bf_int((float)((Ri - a.r0) / a.r0) * 4.0f);
    xv[1] = f2bf_int((float)(wi.x * Ri) * inv_v0);
    xv[2] = f2bf_int((float)(wi.y * Ri) * inv_v0);
    xv[3] = f2bf_int((float)(wi.z * Ri) * inv_v0);
    xv[4] = f2bf_int(rr * rr * rr);
    xv[5] = xv[6] = xv[7] = 0;
    uint4 pk;
    memcpy(&pk, xv, 16);
    *(uint4*)(a.x + i * FN) = pk;
  }
  const uint32_t cnt = a.nl.cnt[i];
  const uint64_t b = a.nl.base[i >> 6] + (uint64_t)(i & 63);
  for (uint32_t k = 0; k < cnt; k++) {
    const uint64_t sl = b + (uint64_t)k * kWave;
    const uint32_t wd = a.nl.word[sl];
    const uint32_t j = wd & kIdxMask;
    const V3 sh = shift_vec(a.g, (int)((wd >> kShiftShift) & 31u));
    const Core cj = a.s.core[j];
    const V3 rj{cj.xr.x + sh.x, cj.xr.y + sh.y, cj.xr.z + sh.z};
    const V3 d = rj - ri;
    const double dist = norm(d);
    const double Rj = cj.xr.w, Rs = Ri + Rj;
    const V3 nn = d / dist;
    const V3 vj{cj.vm.x, cj.vm.y, cj.vm.z}, wj{cj.w.x, cj.w.y, cj.w.z};
    const V3 si = vi + cross(wi, nn * Ri);
    const V3 sj = vj + cross(wj, nn * (-Rj));
    const V3 vr = sj - si;
    const float n0 = (float)nn.x, n1 = (float)nn.y, n2 = (float)nn.z;
    const float o = (float)((Rs - dist) / Rs) * a.ov;
    const float op = fmaxf(o, 0.f), cf = o > 0.f ? 1.f : 0.f;
    const float v0 = (float)vr.x * inv_v0, v1 = (float)vr.y * inv_v0, v2 = (float)vr.z * inv_v0;
    const float vn = cf * ((v0 * n0 + v1 * n1) + v2 * n2);
    uint32_t ev[FE / 2];
    ev[0] = pack_bf16_int(n0, n1);
    ev[1] = pack_bf16_int(n2, o);
    ev[2] = pack_bf16_int(v0, v1);
    ev[3] = pack_bf16_int(v2, (float)((Rj - Ri) / Rs) * 4.0f);
    ev[4] = pack_bf16_int(op * n0, op * n1);
    ev[5] = pack_bf16_int(op * n2, vn * n0);
    ev[6] = pack_bf16_int(vn * n1, vn * n2);
    ev[7] = pack_bf16_int(op, cf);
    uint4 pk[2] = {uint4{ev[0], ev[1], ev[2], ev[3]}, uint4{ev[4], ev[5], ev[6], ev[7]}};
    uint4* dst = (uint4*)(a.e + sl * FE);
    dst[0] = pk[0];
    dst[1] = pk[1];
  }
}

// Reverse entry of every list entry: the slot of j's entry that points back to i with the
// opposite image code.
__global__ __launch_bounds__(256) void k_mirror_all(int64_t n, NbrList nl, uint32_t* __restrict__ mirror) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const uint32_t cnt = nl.cnt[x];
  const uint64_t b = nl.base[x >> 6] + (uint64_t)(x & 63);
  for (uint32_t k = 0; k < cnt; k++) {
    const uint64_t sl = b + (uint64_t)k * kWave;
    const uint32_t wd = nl.word[sl];
    const uint32_t y = wd & kIdxMask;
    const uint32_t want = 26u - ((wd >> kShiftShift) & 31u);
    const uint32_t cy = nl.cnt[y];
    const uint64_t by = nl.base[y >> 6] + (uint64_t)(y & 63);
    uint32_t found = (uint32_t)sl;  // a list is symmetric; keep a valid slot if not found
    for (uint32_t k2 = 0; k2 < cy; k2++) {
      const uint64_t s2 = by + (uint64_t)k2 * kWave;
      const uint32_t w2 = nl.word[s2];
      if ((w2 & kIdxMask) == (uint32_t)x && ((w2 >> kShiftShift) & 31u) == want) {
        found = (uint32_t)s2;
        break;
      }
    }
    mirror[sl] = found;
  }
}

// ------------------------------------------------------------------------------------- encoder
struct EncArgs {
  int64_t n;
  const uint16_t* x;
  const uint16_t* w;  // W_in, bfloat16 [H][FN]
  const uint16_t* b;  // b_in, bfloat16 [H]
  uint16_t* h;
};
// ------------------------------------------------------------------------------------- MFMA GEMM
enum { EPI_BF16 = 0, EPI_UPDATE = 1, EPI_BWD_U = 2, EPI_ADD = 3, EPI_UPDATE_DEC = 4 };

struct GemmArgs {
  int64_t n;
  const uint16_t* a1;
  int k1;
  const uint16_t* a2;
  int k2;
  const uint16_t* b;  // [m][k1 + k2]
  int m;
  const float* bias;
  uint16_t* ybf;    // EPI_BF16 [n][m]
  uint16_t* u_out;  // EPI_UPDATE: pre-activation (may be null)
  uint16_t* h_out;  // EPI_UPDATE: h + relu(U)
  float* dh;        // EPI_BWD_U / EPI_ADD: [n][H] (+=)
  float* dm;        // EPI_BWD_U: [n][H] (=, times 1/8)
  // EPI_UPDATE_DEC: the update of the last layer followed by the decoder on the new hidden state in
  // registers (the state itself is not stored): W_out bfloat16 [FO][H], b_out, outputs [n][FO]
  const uint16_t* dec_w;
  const float* dec_b;
  float* dec_out;
};

// Y[r][c] = sum_k A[r][k] B[c][k] with A = [a1 | a2] (rows = particles), for the fixed shapes of the
// network, computed transposed on MFMA: D[row][particle] with the MFMA A operand = rows of B (the
// weights, staged in LDS) and the MFMA B operand = the particle's row of A (16-byte fragments: lane l
// holds particle l%32 and the 8 consecutive k values 8*(l/32)..+7 of each 16-wide k step).
// Accumulator element v of lane l is MFMA row 8*(v/4) + 4*(l/32) + v%4 of particle column l%32, and
// MFMA rows map to output columns by c(8g + 4h + q) = 16h + 4g + q inside each 32-column tile, so a
// lane holds 16 consecutive output columns of one particle: the epilogue reads and writes 16 bytes.
// The k order is permuted the same way on both operands: k step s = 2t + u takes, in lane half h,
// the 8 inputs 32t + 16h + 8u .. +7, so the fragments a lane loads for k steps 2t and 2t+1 are the
// 16 channels it writes for output tile t (the residual of the update comes from registers).
// Persistent blocks of 8 wavefronts: each block stages all M rows of B once (rows padded by 8
// elements against LDS bank conflicts; 68-70 KB, so two blocks, 16 wavefronts, fit a CU) and walks
// strips of 256 particles (32 per wavefront); a wavefront loads its particles' rows once and forms
// the M columns in passes of 128.
template <int EPI, int M, int K1, int K2>
__global__ __launch_bounds__(512) void k_gemm_p(GemmArgs g) {
  constexpr int K = K1 + K2, KP = K + 8, NT = 4, NPASS = M / 128;
  extern __shared__ uint16_t w_lds[];
  // LDS row t*32 + m holds the weight row of MFMA row m of tile t, i.e. column 32t + c(m): lanes then
  // read consecutive LDS rows, which the 8-element padding spreads over distinct banks
  for (int idx = threadIdx.x; idx < M * K / 8; idx += 512) {
    const int r = idx / (K / 8), c8 = (idx % (K / 8)) * 8;
    const int c = r & 31, m = 8 * ((c >> 2) & 3) + 4 * (c >> 4) + (c & 3);
    *(uint4*)(w_lds + ((r & ~31) + m) * KP + c8) = *(const uint4*)(g.b + (int64_t)r * K + c8);
  }
  // EPI_UPDATE_DEC: the decoder's 32-row tile after the weights (rows FO..31 zero), row stride H + 8
  constexpr int DP = H + 8;
  uint16_t* dec_lds = w_lds + M * KP;
  if (EPI == EPI_UPDATE_DEC) {
    for (int idx = threadIdx.x; idx < 32 * H / 8; idx += 512) {
      const int r = idx / (H / 8), c8 = (idx % (H / 8)) * 8;
      *(uint4*)(dec_lds + r * DP + c8) = r < FO ? *(const uint4*)(g.dec_w + r * H + c8) : uint4{0u, 0u, 0u, 0u};
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int col = lane & 31, half = lane >> 5;
  static_assert(K1 % 32 == 0 && K2 % 32 == 0, "k steps come in pairs");
  static_assert((EPI != EPI_UPDATE && EPI != EPI_UPDATE_DEC) || (M == 128 && K1 == 128),
                "the residual is the first 128 inputs");
  const int64_t ntiles = (g.n + 255) / 256;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t r0 = tile * 256 + w * 32;
    if (r0 >= g.n) continue;
    const int64_t r = r0 + col;  // this lane's particle
    const int64_t ra = r < g.n ? r : g.n - 1;
    // every fragment of the particle rows is requested before the first MFMA
    bf16x8 xf[K / 16];
    const uint16_t* ap1 = g.a1 + ra * K1 + 16 * half;
#pragma unroll
    for (int ks = 0; ks < K1 / 16; ks++) xf[ks] = *(const bf16x8*)(ap1 + 32 * (ks >> 1) + 8 * (ks & 1));
    if (K2 > 0) {
      const uint16_t* ap2 = g.a2 + ra * K2 + 16 * half;
#pragma unroll
      for (int ks = 0; ks < K2 / 16; ks++)
        xf[K1 / 16 + ks] = *(const bf16x8*)(ap2 + 32 * (ks >> 1) + 8 * (ks & 1));
    }
    // keep every fragment load ahead of the first MFMA: left to itself the scheduler sinks each load
    // to its use, and the k steps then wait for memory one after the other (the update product, with
    // twice the fragments, is left to the scheduler: holding all of them costs it occupancy)
    if (EPI != EPI_UPDATE) __builtin_amdgcn_sched_barrier(0);
    // the M output columns in passes of 128 (four 32-column tiles), the fragments held across passes
#pragma unroll
    for (int ps = 0; ps < NPASS; ps++) {
      const int cb = ps * 128;
      floatx16 acc[NT];
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int v = 0; v < 16; v++) acc[t][v] = 0.f;
#pragma unroll
      for (int ks = 0; ks < K / 16; ks++) {
        const int kw = 32 * (ks >> 1) + 16 * half + 8 * (ks & 1);
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const bf16x8 wv = *(const bf16x8*)(w_lds + ((4 * ps + t) * 32 + col) * KP + kw);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv, xf[ks], acc[t], 0, 0, 0);
        }
      }
      floatx16 dacc;
      if (EPI == EPI_UPDATE_DEC) {
#pragma unroll
        for (int v = 0; v < 16; v++) dacc[v] = 0.f;
      }
      // stores only: a lane that left early would run the next pass's (or the decoder's) MFMAs with
      // part of EXEC; the decoder variant runs every lane through the epilogue and guards the stores
#pragma unroll
      for (int t = 0; t < NT && (EPI == EPI_UPDATE_DEC || r < g.n); t++) {
        const int c0 = cb + t * 32 + 16 * half;  // columns c0 .. c0+15
        if (EPI == EPI_BF16) {
          uint32_t o[8];
#pragma unroll
          for (int v = 0; v < 8; v++) o[v] = pack_bf16(acc[t][2 * v], acc[t][2 * v + 1]);
          uint4* dst = (uint4*)(g.ybf + r * M + c0);
          dst[0] = uint4{o[0], o[1], o[2], o[3]};
          dst[1] = uint4{o[4], o[5], o[6], o[7]};
        } else if (EPI == EPI_UPDATE || EPI == EPI_UPDATE_DEC) {
          const uint4 h0 = __builtin_bit_cast(uint4, xf[2 * t]), h1 = __builtin_bit_cast(uint4, xf[2 * t + 1]);
          const uint32_t hw[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
          uint32_t uo[8], ho[8];
#pragma unroll
          for (int v = 0; v < 8; v++) {
            uo[v] = pack_bf16(acc[t][2 * v] + g.bias[c0 + 2 * v], acc[t][2 * v + 1] + g.bias[c0 + 2 * v + 1]);
            const float u0 = bf2f((uint16_t)(uo[v] & 0xffffu)), u1 = bf2f((uint16_t)(uo[v] >> 16));
            ho[v] = pack_bf16(bf2f((uint16_t)(hw[v] & 0xffffu)) + fmaxf(u0, 0.f),
                              bf2f((uint16_t)(hw[v] >> 16)) + fmaxf(u1, 0.f));
          }
          if (EPI == EPI_UPDATE_DEC) {
            // decoder k steps 2t and 2t+1 over the lane's channels 32t + 16h + 0..7 and + 8..15
#pragma unroll
            for (int u = 0; u < 2; u++) {
              const bf16x8 wd = *(const bf16x8*)(dec_lds + col * DP + 32 * t + 16 * half + 8 * u);
              dacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                  wd, __builtin_bit_cast(bf16x8, uint4{ho[4 * u], ho[4 * u + 1], ho[4 * u + 2], ho[4 * u + 3]}), dacc,
                  0, 0, 0);
            }
          } else {
            if (g.u_out) {
              uint4* ud = (uint4*)(g.u_out + r * H + c0);
              ud[0] = uint4{uo[0], uo[1], uo[2], uo[3]};
              ud[1] = uint4{uo[4], uo[5], uo[6], uo[7]};
            }
            uint4* hd = (uint4*)(g.h_out + r * H + c0);
            hd[0] = uint4{ho[0], ho[1], ho[2], ho[3]};
            hd[1] = uint4{ho[4], ho[5], ho[6], ho[7]};
          }
        } else if (EPI == EPI_BWD_U && c0 >= H) {
          float4* dst = (float4*)(g.dm + r * H + (c0 - H));
#pragma unroll
          for (int q = 0; q < 4; q++)
            dst[q] = float4{acc[t][4 * q] * kMsgScale, acc[t][4 * q + 1] * kMsgScale, acc[t][4 * q + 2] * kMsgScale,
                            acc[t][4 * q + 3] * kMsgScale};
        } else {  // EPI_ADD, and the dh half of EPI_BWD_U
          float4* dst = (float4*)(g.dh + r * H + c0);
#pragma unroll
          for (int q = 0; q < 4; q++) {
            float4 cur = dst[q];
            cur.x += acc[t][4 * q];
            cur.y += acc[t][4 * q + 1];
            cur.z += acc[t][4 * q + 2];
            cur.w += acc[t][4 * q + 3];
            dst[q] = cur;
          }
        }
      }
      if (EPI == EPI_UPDATE_DEC && r < g.n) {
        // decoder accumulator element v of lane l is output row 8(v/4) + 4(l/32) + v%4
        if (half == 0) {
#pragma unroll
          for (int q = 0; q < 4; q++) g.dec_out[r * FO + q] = dacc[q] + g.dec_b[q];
        } else {
#pragma unroll
          for (int q = 0; q < 2; q++) g.dec_out[r * FO + 4 + q] = dacc[q] + g.dec_b[4 + q];
        }
      }
    }
  }
}

// Weight gradient Gw[m][k] = sum_r G[r][m] X[r][k] (G [n][mg], X = [x1 | x2] [n][k1+k2], all bf16):
// each block reduces a contiguous range of rows in steps of 16 (the MFMA k dimension), staging the
// 16 rows of G and X in LDS, and writes its mg x K partial sum; k_sum_partials adds the partials.
// mg * K = 32768: 32 tiles of 32x32, 8 per wavefront.
struct WgradArgs {
  int64_t n;
  const uint16_t* gm;
  int mg;
  const uint16_t* x1;
  int k1;
  const uint16_t* x2;
  int k2;
  int64_t rows_per_block;
  float* partial;  // [gridDim.x][mg][K]
};
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
  __shared__ uint16_t sg[16 * 256];
  __shared__ uint16_t sx[16 * 256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int K = a.k1 + a.k2;
  const int tk = K / 32;
  floatx16 acc[8];
#pragma unroll
  for (int t = 0; t < 8; t++)
    for (int v = 0; v < 16; v++) acc[t][v] = 0.f;
  const int64_t rbeg = (int64_t)blockIdx.x * a.rows_per_block;
  const int64_t rend = std::min<int64_t>(a.n, rbeg + a.rows_per_block);
  for (int64_t r0 = rbeg; r0 < rend; r0 += 16) {
    __syncthreads();
    // stage 16 rows: G row-major [16][mg], X [16][K]; zero beyond the range
    for (int idx = threadIdx.x; idx < 16 * a.mg / 8; idx += 256) {
      const int rr = idx / (a.mg / 8), cc = (idx % (a.mg / 8)) * 8;
      uint4 v = {0, 0, 0, 0};
      if (r0 + rr < rend) v = *(const uint4*)(a.gm + (r0 + rr) * a.mg + cc);
      *(uint4*)(sg + rr * a.mg + cc) = v;
    }
    for (int idx = threadIdx.x; idx < 16 * K / 8; idx += 256) {
      const int rr = idx / (K / 8), cc = (idx % (K / 8)) * 8;
      uint4 v = {0, 0, 0, 0};
      if (r0 + rr < rend) {
        if (cc < a.k1) v = *(const uint4*)(a.x1 + (r0 + rr) * a.k1 + cc);
        else v = *(const uint4*)(a.x2 + (r0 + rr) * a.k2 + (cc - a.k1));
      }
      *(uint4*)(sx + rr * K + cc) = v;
    }
    __syncthreads();
    const int rk = 8 * (lane >> 5);
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int tile = w * 8 + t;
      const int ti = tile / tk, tj = tile % tk;
      const int mrow = ti * 32 + (lane & 31), kcol = tj * 32 + (lane & 31);
      bf16x8 av, bv;
      uint16_t ta[8], tb[8];
#pragma unroll
      for (int s = 0; s < 8; s++) {
        ta[s] = sg[(rk + s) * a.mg + mrow];
        tb[s] = sx[(rk + s) * K + kcol];
      }
      memcpy(&av, ta, 16);
      memcpy(&bv, tb, 16);
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[t], 0, 0, 0);
    }
  }
  float* out = a.partial + (int64_t)blockIdx.x * a.mg * K;
#pragma unroll
  for (int t = 0; t < 8; t++) {
    const int tile = w * 8 + t;
    const int ti = tile / tk, tj = tile % tk;
    for (int v = 0; v < 16; v++) {
      const int mrow = ti * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
      const int kcol = tj * 32 + (lane & 31);
      out[(int64_t)mrow * K + kcol] = acc[t][v];
    }
  }
}

// out[k] = sum over p < np of partial[p][k], in order of p
__global__ __launch_bounds__(256) void k_sum_partials(const float* __restrict__ partial, int np, int64_t len,
                                                      float* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= len) return;
  float s = 0.f;
  for (int p = 0; p < np; p++) s += partial[(int64_t)p * len + k];
  out[k] = s;
}

// column sums of a bf16 [n][m] matrix (bias gradients): block b sums rows b, b+B, ... per column
__global__ __launch_bounds__(256) void k_colsum(const uint16_t* __restrict__ gm, int64_t n, int m,
                                                float* __restrict__ partial) {
  for (int c = threadIdx.x; c < m; c += blockDim.x) {
    float s = 0.f;
    for (int64_t r = blockIdx.x; r < n; r += gridDim.x) s += bf2f(gm[r * m + c]);
    partial[(int64_t)blockIdx.x * m + c] = s;
  }
}

// ------------------------------------------------------------------------------------- messages
struct MsgArgs {
  int64_t n;
  NbrList nl;
  const uint16_t* e;
  const uint16_t* pq;
  const float* wE;  // [H][FE] rounded
  const float* bE;
  uint16_t* m;       // forward output [n][H]
  const float* dm;   // backward input [n][H]
  const uint32_t* mirror;
  uint16_t* dpq;     // backward output [n][2H]
  float* part_w;     // backward partials [blocks][H][FE]
  float* part_b;     // [blocks][H]
};

__device__ inline void load_e(const uint16_t* p, float ev[FE]) {
  uint4 pk[2];
  pk[0] = ((const uint4*)p)[0];
  pk[1] = ((const uint4*)p)[1];
  uint16_t t[FE];
  memcpy(t, pk, 32);
#pragma unroll
  for (int k = 0; k < FE; k++) ev[k] = bf2f(t[k]);
}
__device__ inline float edge_z(float P, float Q, const float ev[FE], const float w[FE], float b) {
  float E = 0.f;
#pragma unroll
  for (int k = 0; k < FE; k++) E = fmaf(w[k], ev[k], E);
  return ((P + Q) + E) + b;
}
__device__ inline void ld2(const uint16_t* p, float& a, float& b) {
  const uint32_t v = *(const uint32_t*)p;
  a = bf2f((uint16_t)(v & 0xffffu));
  b = bf2f((uint16_t)(v >> 16));
}

// Message pass on MFMA: one block of four wavefronts per 32 consecutive particles (half of a
// 64-particle list group), each wavefront for 32 of the 128 channels. Step k takes neighbour rank k
// of each of the 32 particles, i.e. 32 consecutive list slots, and computes D[row][edge] =
// sum_f W_E[c(row)][f] e[edge][f] + (P_i + b_E)[c(row)] + Q_j[c(row)] with three 32x32x16 MFMAs:
// A = W_E rows and B = the edges' 16 inputs, C = P_i + b_E; then A = identity rows and B = the 32
// channels of Q_j, 16 per MFMA (lane half h holds channels 8h..8h+7 of each half), so that Q_j is
// added on the matrix unit; relu(z) is summed over k in the accumulator layout. Accumulator element
// v of lane l is row 8*(v/4) + 4*(l/32) + v%4 of edge column l%32; rows are mapped to channels by
// c(8g + 4h + q) = 16h + 4g + q inside the 32-channel tile, so that lane half h holds the 16
// consecutive channels 16h..16h+15 (16-byte loads of P, 16-byte stores of m). The list words and edge inputs of the block's 32 slots are shared by its four
// wavefronts through L1. The list word, edge inputs and Q rows of step k+1 are requested before
// step k is computed, the list word of step k+2 before that.
struct MsgStep {
  bool ok;
  bf16x8 eb;
  uint4 q[2];
};
__device__ inline void msg_load(const MsgArgs& a, uint64_t sl, bool ok, uint32_t wd, int half, int ch0, uint32_t self,
                                MsgStep& st) {
  st.ok = ok;
  (void)self;
  if (ok) {
    st.eb = *(const bf16x8*)(a.e + sl * FE + 8 * half);
    const uint16_t* qrow = a.pq + (int64_t)(wd & kIdxMask) * H2 + H + ch0 + 8 * half;
    st.q[0] = *(const uint4*)(qrow);
    st.q[1] = *(const uint4*)(qrow + 16);
  } else {
    // padding slot of the list group: nothing is gathered (its column of the MFMA is not summed)
#pragma unroll
    for (int q = 0; q < 8; q++) st.eb[q] = (__bf16)0.0f;
    st.q[0] = uint4{0u, 0u, 0u, 0u};
    st.q[1] = uint4{0u, 0u, 0u, 0u};
  }
}
__device__ inline float bf_lo(uint32_t v) { return bf2f((uint16_t)(v & 0xffffu)); }
__device__ inline float bf_hi(uint32_t v) { return bf2f((uint16_t)(v >> 16)); }

__global__ __launch_bounds__(256) void k_msg_fwd_mfma(MsgArgs a, const uint16_t* __restrict__ wEb) {
  const int lane = threadIdx.x & 63;
  // XCD-aware block order: consecutive particle groups on one XCD share its L2 for the Q gathers
  const int64_t i0 = xcd_block(blockIdx.x, gridDim.x) * 32;
  const int ch0 = (int)(threadIdx.x >> 6) * 32;
  if (i0 >= a.n) return;
  const int col = lane & 31, half = lane >> 5;
  const int64_t i = i0 + col;
  const bool vi = i < a.n;
  const int64_t ic = vi ? i : a.n - 1;
  const int64_t grp = i0 >> 6;
  const uint64_t base = a.nl.base[grp] + (uint64_t)(ic & 63);
  const uint32_t cnt = vi ? a.nl.cnt[ic] : 0u;
  const uint32_t kmax = a.nl.kmax[grp];
  // A rows: MFMA row `col` of the tile is channel c(col)
  const int crow = 16 * ((col >> 2) & 1) + 4 * (col >> 3) + (col & 3);
  const bf16x8 wa = *(const bf16x8*)(wEb + (ch0 + crow) * FE + 8 * half);
  // identity rows: MFMA row `col` (channel crow) picks input channel crow from the Q fragments
  uint32_t ia[4], ib[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int k0 = 8 * half + 2 * q;
    ia[q] = (crow == k0 ? 0x3f80u : 0u) | (crow == k0 + 1 ? 0x3f800000u : 0u);
    ib[q] = (crow == 16 + k0 ? 0x3f80u : 0u) | (crow == 16 + k0 + 1 ? 0x3f800000u : 0u);
  }
  const bf16x8 wia = __builtin_bit_cast(bf16x8, uint4{ia[0], ia[1], ia[2], ia[3]});
  const bf16x8 wib = __builtin_bit_cast(bf16x8, uint4{ib[0], ib[1], ib[2], ib[3]});
  const int c0 = ch0 + 16 * half;
  floatx16 cinit, msum;
  {
    const uint4 p0 = *(const uint4*)(a.pq + ic * H2 + c0);
    const uint4 p1 = *(const uint4*)(a.pq + ic * H2 + c0 + 8);
    const uint32_t pw[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
#pragma unroll
    for (int v = 0; v < 8; v++) {
      cinit[2 * v] = bf_lo(pw[v]) + a.bE[c0 + 2 * v];
      cinit[2 * v + 1] = bf_hi(pw[v]) + a.bE[c0 + 2 * v + 1];
    }
#pragma unroll
    for (int v = 0; v < 16; v++) msum[v] = 0.f;
  }
  MsgStep cur, nxt;
  uint32_t wnext = 0;
  if (kmax > 0) {
    const uint32_t w0 = cnt > 0 ? a.nl.word[base] : 0u;
    msg_load(a, base, cnt > 0, w0, half, ch0, (uint32_t)ic, cur);
    if (kmax > 1 && cnt > 1) wnext = a.nl.word[base + kWave];
  }
  for (uint32_t k = 0; k < kmax; k++) {
    if (k + 1 < kmax) {
      const uint64_t sl1 = base + (uint64_t)(k + 1) * kWave;
      msg_load(a, sl1, k + 1 < cnt, wnext, half, ch0, (uint32_t)ic, nxt);
      wnext = (k + 2 < kmax && k + 2 < cnt) ? a.nl.word[sl1 + kWave] : 0u;
    }
    floatx16 d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, cur.eb, cinit, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wia, __builtin_bit_cast(bf16x8, cur.q[0]), d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wib, __builtin_bit_cast(bf16x8, cur.q[1]), d, 0, 0, 0);
    if (cur.ok) {
#pragma unroll
      for (int v = 0; v < 16; v++) msum[v] += fmaxf(d[v], 0.f);
    }
    cur = nxt;
  }
  if (!vi) return;
  uint32_t o[8];
#pragma unroll
  for (int v = 0; v < 8; v++)
    o[v] = pack_bf16(msum[2 * v] * kMsgScale, msum[2 * v + 1] * kMsgScale);
  *(uint4*)(a.m + i * H + c0) = uint4{o[0], o[1], o[2], o[3]};
  *(uint4*)(a.m + i * H + c0 + 8) = uint4{o[4], o[5], o[6], o[7]};
}

// backward of the messages: per particle j, dP_j = sum over own entries j->i of dm_j (z_ji > 0),
// dQ_j = sum over the reverse entries i->j of dm_i (z_ij > 0); W_E and b_E gradients over own entries
__global__ __launch_bounds__(256) void k_msg_bwd(MsgArgs a) {
  __shared__ float red[4][H * (FE + 1)];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = 2 * lane;
  float w0[FE], w1[FE];
#pragma unroll
  for (int k = 0; k < FE; k++) {
    w0[k] = a.wE[c * FE + k];
    w1[k] = a.wE[(c + 1) * FE + k];
  }
  const float b0 = a.bE[c], b1 = a.bE[c + 1];
  float gw0[FE], gw1[FE], gb0 = 0.f, gb1 = 0.f;
#pragma unroll
  for (int k = 0; k < FE; k++) gw0[k] = gw1[k] = 0.f;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < a.n; j += nw) {
    float Pj0, Pj1, Qj0, Qj1;
    ld2(a.pq + j * H2 + c, Pj0, Pj1);
    ld2(a.pq + j * H2 + H + c, Qj0, Qj1);
    const float dmj0 = a.dm[j * H + c], dmj1 = a.dm[j * H + c + 1];
    const uint32_t cnt = a.nl.cnt[j];
    const uint64_t base = a.nl.base[j >> 6] + (uint64_t)(j & 63);
    float dP0 = 0.f, dP1 = 0.f, dQ0 = 0.f, dQ1 = 0.f;
    for (uint32_t k = 0; k < cnt; k++) {
      const uint64_t sl = base + (uint64_t)k * kWave;
      const uint32_t i = a.nl.word[sl] & kIdxMask;
      float Pi0, Pi1, Qi0, Qi1;
      ld2(a.pq + (int64_t)i * H2 + c, Pi0, Pi1);
      ld2(a.pq + (int64_t)i * H2 + H + c, Qi0, Qi1);
      // own entry j -> i
      float ev[FE];
      load_e(a.e + sl * FE, ev);
      const float g0 = edge_z(Pj0, Qi0, ev, w0, b0) > 0.f ? dmj0 : 0.f;
      const float g1 = edge_z(Pj1, Qi1, ev, w1, b1) > 0.f ? dmj1 : 0.f;
      dP0 += g0;
      dP1 += g1;
      gb0 += g0;
      gb1 += g1;
#pragma unroll
      for (int q = 0; q < FE; q++) {
        gw0[q] = fmaf(g0, ev[q], gw0[q]);
        gw1[q] = fmaf(g1, ev[q], gw1[q]);
      }
      // reverse entry i -> j
      const uint64_t ms = a.mirror[sl];
      float er[FE];
      load_e(a.e + ms * FE, er);
      const float dmi0 = a.dm[(int64_t)i * H + c], dmi1 = a.dm[(int64_t)i * H + c + 1];
      dQ0 += edge_z(Pi0, Qj0, er, w0, b0) > 0.f ? dmi0 : 0.f;
      dQ1 += edge_z(Pi1, Qj1, er, w1, b1) > 0.f ? dmi1 : 0.f;
    }
    *(uint32_t*)(a.dpq + j * H2 + c) = pack_bf16(dP0, dP1);
    *(uint32_t*)(a.dpq + j * H2 + H + c) = pack_bf16(dQ0, dQ1);
  }
  // block reduction of the W_E / b_E gradients in wavefront order
#pragma unroll
  for (int k = 0; k < FE; k++) {
    red[wv][c * (FE + 1) + k] = gw0[k];
    red[wv][(c + 1) * (FE + 1) + k] = gw1[k];
  }
  red[wv][c * (FE + 1) + FE] = gb0;
  red[wv][(c + 1) * (FE + 1) + FE] = gb1;
  __syncthreads();
  for (int idx = threadIdx.x; idx < H * (FE + 1); idx += blockDim.x) {
    const float s = ((red[0][idx] + red[1][idx]) + red[2][idx]) + red[3][idx];
    const int ch = idx / (FE + 1), k = idx % (FE + 1);
    if (k < FE) a.part_w[(int64_t)blockIdx.x * H * FE + ch * FE + k] = s;
    else a.part_b[(int64_t)blockIdx.x * H + ch] = s;
  }
}

// ------------------------------------------------------------------------------------- decoder
struct DecArgs {
  int64_t n;
  const uint16_t* h;
  const float* w;  // W_out rounded [FO][H]
  const uint16_t* wb;  // W_out, bfloat16 [FO][H] (forward on MFMA)
  const float* b;
  float* out;           // [n][FO]
  const float* target;  // training: [n][FO]
  float* dh;            // [n][H]
  float* part;          // [blocks][FO*H + FO + 1] (dW_out, db_out, loss)
};
__device__ inline float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_decode(DecArgs a) {
  __shared__ float red[4][FO * H + FO + 1];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = 2 * lane;
  float wo0[FO], wo1[FO];
#pragma unroll
  for (int o = 0; o < FO; o++) {
    wo0[o] = a.w[o * H + c];
    wo1[o] = a.w[o * H + c + 1];
  }
  float gw0[FO], gw1[FO], gb[FO], gl = 0.f;
#pragma unroll
  for (int o = 0; o < FO; o++) gw0[o] = gw1[o] = gb[o] = 0.f;
  const float scale = TRAIN ? 2.0f / (float)(a.n * FO) : 0.f;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < a.n; i += nw) {
    float h0, h1;
    ld2(a.h + i * H + c, h0, h1);
    float y[FO];
#pragma unroll
    for (int o = 0; o < FO; o++) y[o] = wave_sum(fmaf(wo1[o], h1, wo0[o] * h0)) + a.b[o];
    if (lane < FO) {
      float yo = y[0];
#pragma unroll
      for (int o = 1; o < FO; o++)
        if (lane == o) yo = y[o];
      a.out[i * FO + lane] = yo;
    }
    if (TRAIN) {
      float d0 = 0.f, d1 = 0.f;
#pragma unroll
      for (int o = 0; o < FO; o++) {
        const float diff = y[o] - a.target[i * FO + o];
        const float dout = scale * diff;
        if (lane == 0) {
          gl = fmaf(diff, diff, gl);
          gb[o] += dout;
        }
        d0 = fmaf(wo0[o], dout, d0);
        d1 = fmaf(wo1[o], dout, d1);
        gw0[o] = fmaf(dout, h0, gw0[o]);
        gw1[o] = fmaf(dout, h1, gw1[o]);
      }
      a.dh[i * H + c] = d0;
      a.dh[i * H + c + 1] = d1;
    }
  }
  if (TRAIN) {
#pragma unroll
    for (int o = 0; o < FO; o++) {
      red[wv][o * H + c] = gw0[o];
      red[wv][o * H + c + 1] = gw1[o];
    }
    if (lane == 0) {
#pragma unroll
      for (int o = 0; o < FO; o++) red[wv][FO * H + o] = gb[o];
      red[wv][FO * H + FO] = gl;
    }
    __syncthreads();
    const int len = FO * H + FO + 1;
    for (int idx = threadIdx.x; idx < len; idx += blockDim.x)
      a.part[(int64_t)blockIdx.x * len + idx] = ((red[0][idx] + red[1][idx]) + red[2][idx]) + red[3][idx];
  }
}


// Encoder on MFMA: h0 = relu(x W_in^T + b_in), one wavefront per 32 particles. The B operand is a
// particle's 8 inputs (lane half 0) and a constant 1 at k = 8 (lane half 1), whose weight is the
// bias, so the bias is part of the product. MFMA rows map to channels as in the dense products:
// lane half h holds channels 32t + 16h .. +15 of its particle, stored with 16-byte writes.
__global__ __launch_bounds__(256) void k_encode_mfma(EncArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t p0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
  if (p0 >= a.n) return;
  const int col = lane & 31, half = lane >> 5;
  const int crow = 16 * ((col >> 2) & 1) + 4 * (col >> 3) + (col & 3);
  const int64_t p = p0 + col;
  const int64_t pc = p < a.n ? p : a.n - 1;
  const uint4 xin = *(const uint4*)(a.x + pc * FN);
  const bf16x8 xb = __builtin_bit_cast(bf16x8, half ? uint4{0x3f80u, 0u, 0u, 0u} : xin);
  bf16x8 wa[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int ch = 32 * t + crow;
    const uint4 wrow = *(const uint4*)(a.w + ch * FN);
    wa[t] = __builtin_bit_cast(bf16x8, half ? uint4{(uint32_t)a.b[ch], 0u, 0u, 0u} : wrow);
  }
  floatx16 zero;
#pragma unroll
  for (int v = 0; v < 16; v++) zero[v] = 0.f;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const floatx16 acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[t], xb, zero, 0, 0, 0);
    uint32_t o[8];
#pragma unroll
    for (int v = 0; v < 8; v++)
      o[v] = pack_bf16(fmaxf(acc[2 * v], 0.f), fmaxf(acc[2 * v + 1], 0.f));
    if (p < a.n) {
      uint4* dst = (uint4*)(a.h + p * H + 32 * t + 16 * half);
      dst[0] = uint4{o[0], o[1], o[2], o[3]};
      dst[1] = uint4{o[4], o[5], o[6], o[7]};
    }
  }
}

// Decoder on MFMA (forward): out = h W_out^T + b_out, one wavefront per 32 particles. The FO = 6
// outputs are rows 0..5 of one 32-row tile (rows 6..31 have zero weights); k step s takes channels
// 16s + 8h .. +7 in lane half h. Accumulator element v of lane l is output row 8(v/4) + 4(l/32) + v%4.
__global__ __launch_bounds__(256) void k_decode_mfma(DecArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t p0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
  if (p0 >= a.n) return;
  const int col = lane & 31, half = lane >> 5;
  const int64_t p = p0 + col;
  const int64_t pc = p < a.n ? p : a.n - 1;
  bf16x8 hf[H / 16], wf[H / 16];
  const int wr = col < FO ? col : 0;  // rows past FO read row 0 and are zeroed
#pragma unroll
  for (int s = 0; s < H / 16; s++) {
    hf[s] = *(const bf16x8*)(a.h + pc * H + 16 * s + 8 * half);
    const uint4 w = *(const uint4*)(a.wb + wr * H + 16 * s + 8 * half);
    wf[s] = __builtin_bit_cast(bf16x8, col < FO ? w : uint4{0u, 0u, 0u, 0u});
  }
  __builtin_amdgcn_sched_barrier(0);
  floatx16 acc;
#pragma unroll
  for (int v = 0; v < 16; v++) acc[v] = 0.f;
#pragma unroll
  for (int s = 0; s < H / 16; s++) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wf[s], hf[s], acc, 0, 0, 0);
  if (p >= a.n) return;
  if (half == 0) {
#pragma unroll
    for (int r = 0; r < 4; r++) a.out[p * FO + r] = acc[r] + a.b[r];
  } else {
#pragma unroll
    for (int r = 0; r < 2; r++) a.out[p * FO + 4 + r] = acc[r] + a.b[4 + r];
  }
}

// ------------------------------------------------------------------------------------- elementwise
// du = bf16(dh * (U > 0))
__global__ __launch_bounds__(256) void k_relu_grad(int64_t len, const float* __restrict__ dh,
                                                   const uint16_t* __restrict__ u, uint16_t* __restrict__ du) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= len) return;
  du[k] = f2bf(bf2f(u[k]) > 0.f ? dh[k] : 0.f);
}

// encoder backward: dpre = dh * (h0 > 0); W_in and b_in gradients, per-block partials
__global__ __launch_bounds__(256) void k_encode_bwd(int64_t n, const uint16_t* __restrict__ x,
                                                    const uint16_t* __restrict__ h0, const float* __restrict__ dh,
                                                    float* __restrict__ part) {
  __shared__ float red[4][H * (FN + 1)];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = 2 * lane;
  float g0[FN + 1], g1[FN + 1];
#pragma unroll
  for (int k = 0; k <= FN; k++) g0[k] = g1[k] = 0.f;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n; i += nw) {
    float hv0, hv1;
    ld2(h0 + i * H + c, hv0, hv1);
    const float d0 = hv0 > 0.f ? dh[i * H + c] : 0.f;
    const float d1 = hv1 > 0.f ? dh[i * H + c + 1] : 0.f;
#pragma unroll
    for (int k = 0; k < FN; k++) {
      const float xv = bf2f(x[i * FN + k]);
      g0[k] = fmaf(d0, xv, g0[k]);
      g1[k] = fmaf(d1, xv, g1[k]);
    }
    g0[FN] += d0;
    g1[FN] += d1;
  }
#pragma unroll
  for (int k = 0; k <= FN; k++) {
    red[wv][c * (FN + 1) + k] = g0[k];
    red[wv][(c + 1) * (FN + 1) + k] = g1[k];
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < H * (FN + 1); idx += blockDim.x)
    part[(int64_t)blockIdx.x * H * (FN + 1) + idx] = ((red[0][idx] + red[1][idx]) + red[2][idx]) + red[3][idx];
}

// scatter a [blocks][H][FN+1] / [H][FE+1]-style interleaved reduction into weight and bias gradients
__global__ __launch_bounds__(256) void k_split_wb(const float* __restrict__ sum, int rows, int cols,
                                                  float* __restrict__ gw, float* __restrict__ gb) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * (cols + 1)) return;
  const int r = idx / (cols + 1), k = idx % (cols + 1);
  if (k < cols) gw[r * cols + k] = sum[idx];
  else gb[r] = sum[idx];
}

// Adam, then the bf16 working copy and its fp32 image
__global__ __launch_bounds__(256) void k_adam(int64_t np, float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m1, float* __restrict__ m2, float lr, float b1,
                                              float b2, float eps, float c1, float c2, uint16_t* __restrict__ wbf,
                                              float* __restrict__ wr) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const float gk = g[k];
  const float a = b1 * m1[k] + (1.f - b1) * gk;
  const float b = b2 * m2[k] + (1.f - b2) * gk * gk;
  m1[k] = a;
  m2[k] = b;
  const float pk = p[k] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  p[k] = pk;
  const uint16_t q = f2bf(pk);
  wbf[k] = q;
  wr[k] = bf2f(q);
}
__global__ __launch_bounds__(256) void k_round_params(int64_t np, const float* __restrict__ p,
                                                      uint16_t* __restrict__ wbf, float* __restrict__ wr) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const uint16_t q = f2bf(p[k]);
  wbf[k] = q;
  wr[k] = bf2f(q);
}
// out[c][r] = in[r][c] for an rows x cols bf16 matrix
__global__ __launch_bounds__(256) void k_transpose(const uint16_t* __restrict__ in, int rows, int cols,
                                                   uint16_t* __restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * cols) return;
  const int r = idx / cols, c = idx % cols;
  out[(int64_t)c * rows + r] = in[idx];
}

// ------------------------------------------------------------------------------------- surrogate
// velocity / angular-velocity snapshot
__global__ __launch_bounds__(256) void k_snapshot(int64_t n, PState s, double* __restrict__ snap) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Core c = s.core[i];
  double* o = snap + i * 6;
  o[0] = c.vm.x; o[1] = c.vm.y; o[2] = c.vm.z;
  o[3] = c.w.x; o[4] = c.w.y; o[5] = c.w.z;
}
// targets of a K-step window: the mean contact force and torque in units of m0 a0 and m0 a0 r0,
// F = m ((v - v0)/T - g), tau = I (w - w0)/T (sphere inertia is isotropic)
__global__ __launch_bounds__(256) void k_targets(int64_t n, PState s, const double* __restrict__ snap, double T,
                                                 V3 g, double fs, double ts, float* __restrict__ t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Core c = s.core[i];
  const double* o = snap + i * 6;
  const double m = c.vm.w;
  float* q = t + i * FO;
  q[0] = (float)(m * ((c.vm.x - o[0]) / T - g.x) / fs);
  q[1] = (float)(m * ((c.vm.y - o[1]) / T - g.y) / fs);
  q[2] = (float)(m * ((c.vm.z - o[2]) / T - g.z) / fs);
  q[3] = (float)(s.ix[i] * (c.w.x - o[3]) / T / ts);
  q[4] = (float)(s.iy[i] * (c.w.y - o[4]) / T / ts);
  q[5] = (float)(s.iz[i] * (c.w.z - o[5]) / T / ts);
}
// one surrogate step of length T: a = g + fs out / m, alpha = ts out / I;
// x += v T + a T^2/2, v += a T, w += alpha T, Q <- normalize(Q + dot(Q, w) T) (push_to_quaternion form).
// When check != 0, the predicted positions and the free-flight positions (gravity only) are stored
// by particle id for the drift check.
__global__ __launch_bounds__(256) void k_sur_advance(int64_t n, PState s, const float* __restrict__ out, double T,
                                                     V3 g, double fs, double ts, int apply, int check,
                                                     double* __restrict__ xs, double* __restrict__ xf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Core c = s.core[i];
  const float* o = out + i * FO;
  const double m = c.vm.w;
  const V3 a{g.x + fs * o[0] / m, g.y + fs * o[1] / m, g.z + fs * o[2] / m};
  const V3 al{ts * o[3] / s.ix[i], ts * o[4] / s.iy[i], ts * o[5] / s.iz[i]};
  const double T2 = 0.5 * T * T;
  const V3 x{c.xr.x + c.vm.x * T + a.x * T2, c.xr.y + c.vm.y * T + a.y * T2, c.xr.z + c.vm.z * T + a.z * T2};
  if (check) {
    const uint64_t id = s.id[i];
    xs[id * 3 + 0] = x.x;
    xs[id * 3 + 1] = x.y;
    xs[id * 3 + 2] = x.z;
    xf[id * 3 + 0] = c.xr.x + c.vm.x * T + g.x * T2;
    xf[id * 3 + 1] = c.xr.y + c.vm.y * T + g.y * T2;
    xf[id * 3 + 2] = c.xr.z + c.vm.z * T + g.z * T2;
  }
  if (!apply) return;
  const V3 w{c.w.x, c.w.y, c.w.z};
  // dot(Q, w): quaternion derivative 0.5 * (0, w) * Q (push_to_quaternion.hpp:31-37 form)
  const double qw = c.q.x, qx = c.q.y, qy = c.q.z, qz = c.q.w;
  const double dw = -0.5 * (w.x * qx + w.y * qy + w.z * qz);
  const double dx = 0.5 * (w.x * qw + w.y * qz - w.z * qy);
  const double dy = 0.5 * (-w.x * qz + w.y * qw + w.z * qx);
  const double dz = 0.5 * (w.x * qy - w.y * qx + w.z * qw);
  double nw_ = qw + dw * T, nx = qx + dx * T, ny = qy + dy * T, nz = qz + dz * T;
  const double inv = 1.0 / sqrt(nw_ * nw_ + nx * nx + ny * ny + nz * nz);
  c.q = {nw_ * inv, nx * inv, ny * inv, nz * inv};
  c.xr.x = x.x;
  c.xr.y = x.y;
  c.xr.z = x.z;
  c.vm.x += a.x * T;
  c.vm.y += a.y * T;
  c.vm.z += a.z * T;
  c.w.x += al.x * T;
  c.w.y += al.y * T;
  c.w.z += al.z * T;
  s.core[i] = c;
}
// largest |x_dem - x_ref[id]| / R per block (two references)
__global__ __launch_bounds__(256) void k_drift(int64_t n, PState s, const double* __restrict__ xs,
                                               const double* __restrict__ xf, double* __restrict__ part) {
  // per block: largest and summed squared |x - x_ref| / R against the surrogate (xs) and the free
  // flight (xf) positions
  __shared__ double r[4][256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double e1 = 0, e2 = 0;
  if (i < n) {
    const Core c = s.core[i];
    const uint64_t id = s.id[i];
    const V3 x{c.xr.x, c.xr.y, c.xr.z};
    const V3 d1{x.x - xs[id * 3], x.y - xs[id * 3 + 1], x.z - xs[id * 3 + 2]};
    const V3 d2{x.x - xf[id * 3], x.y - xf[id * 3 + 1], x.z - xf[id * 3 + 2]};
    e1 = norm(d1) / c.xr.w;
    e2 = norm(d2) / c.xr.w;
  }
  r[0][threadIdx.x] = e1;
  r[1][threadIdx.x] = e2;
  r[2][threadIdx.x] = e1 * e1;
  r[3][threadIdx.x] = e2 * e2;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      r[0][threadIdx.x] = fmax(r[0][threadIdx.x], r[0][threadIdx.x + st]);
      r[1][threadIdx.x] = fmax(r[1][threadIdx.x], r[1][threadIdx.x + st]);
      r[2][threadIdx.x] += r[2][threadIdx.x + st];
      r[3][threadIdx.x] += r[3][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = r[threadIdx.x][0];
}

}  // namespace gnn
}  // namespace xdem

using namespace xdem;
using namespace xdem::gnn;

struct xdem_gnn {
  xdem_ctx* c = nullptr;
  xdem_gnn_config cfg{};
  Off off{};
  float* p = nullptr;   // fp32 master parameters
  float* g = nullptr;   // gradients
  float* m1 = nullptr;  // Adam moments
  float* m2 = nullptr;
  uint16_t* wbf = nullptr;   // bf16 copy (MFMA operands)
  float* wr = nullptr;       // fp32 image of the bf16 copy (VALU kernels)
  uint16_t* wT = nullptr;    // per layer: W_AB^T [H][2H], W_U^T [2H][H]
  int64_t adam_t = 0;
  // activations
  int64_t cap_n = 0, cap_slots = 0;
  uint16_t* x = nullptr;
  uint16_t* e = nullptr;
  uint16_t* h[L + 1] = {};
  uint16_t* pq[L] = {};
  uint16_t* m[L] = {};
  uint16_t* u[L] = {};
  float* out = nullptr;
  float* target = nullptr;
  float* dh = nullptr;
  float* dm = nullptr;
  uint16_t* du = nullptr;
  uint16_t* dpq = nullptr;
  uint32_t* mirror = nullptr;
  float* partial = nullptr;  // partial sums of the reductions (sized once, never reallocated)
  int64_t cap_partial = 0;
  float* dec_sum = nullptr;  // decoder reduction: dW_out, db_out, loss
  double* snap = nullptr;
  double* xs = nullptr;
  double* xf = nullptr;
  double* dpart = nullptr;
  double* h_dpart = nullptr;
  float* h_loss = nullptr;
  bool fwd_valid = false;
  int64_t fwd_n = 0;
};

namespace {

template <typename T>
int galloc(xdem_ctx* c, T** p, int64_t count) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  XDEM_HIP(c, hipMalloc((void**)p, sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
  return XDEM_OK;
}

int ensure_buffers(xdem_gnn* G) {
  xdem_ctx* c = G->c;
  const int64_t n = c->n;
  const NbrList& nl = c->nl[c->nl_cur];
  int rc;
  if (nl.slots > G->cap_slots) {
    const int64_t cap = (int64_t)(nl.slots * 1.25) + 4096;
    if ((rc = galloc(c, &G->e, cap * FE))) return rc;
    if (G->cfg.train && (rc = galloc(c, &G->mirror, cap))) return rc;
    G->cap_slots = cap;
  }
  if (n > G->cap_n) {
    const int64_t cap = (int64_t)(n * 1.1) + 256;
    const int nh = G->cfg.train ? L + 1 : 2;
    const int nl_ = G->cfg.train ? L : 1;
    if ((rc = galloc(c, &G->x, cap * FN))) return rc;
    for (int k = 0; k < nh; k++)
      if ((rc = galloc(c, &G->h[k], cap * H))) return rc;
    for (int k = 0; k < nl_; k++) {
      if ((rc = galloc(c, &G->pq[k], cap * H2))) return rc;
      if ((rc = galloc(c, &G->m[k], cap * H))) return rc;
      if (G->cfg.train && (rc = galloc(c, &G->u[k], cap * H))) return rc;
    }
    if ((rc = galloc(c, &G->out, cap * FO))) return rc;
    if (G->cfg.train) {
      if ((rc = galloc(c, &G->target, cap * FO)) || (rc = galloc(c, &G->dh, cap * H)) ||
          (rc = galloc(c, &G->dm, cap * H)) || (rc = galloc(c, &G->du, cap * H)) ||
          (rc = galloc(c, &G->dpq, cap * H2)) || (rc = galloc(c, &G->snap, cap * 6)))
        return rc;
    }
    if ((rc = galloc(c, &G->xs, cap * 3)) || (rc = galloc(c, &G->xf, cap * 3))) return rc;
    G->cap_n = cap;
  }
  return XDEM_OK;
}

inline int nb(int64_t n, int bs) { return (int)std::max<int64_t>(1, (n + bs - 1) / bs); }
inline double force_scale(const xdem_gnn* G) { return G->cfg.m0 * G->cfg.a0; }
inline double torque_scale(const xdem_gnn* G) { return G->cfg.m0 * G->cfg.a0 * G->cfg.r0; }

int refresh_weights(xdem_gnn* G) {
  xdem_ctx* c = G->c;
  for (int l = 0; l < L; l++) {
    uint16_t* tab = G->wT + (int64_t)l * 2 * H * H2;
    hipLaunchKernelGGL(k_transpose, dim3(nb((int64_t)H2 * H, 256)), dim3(256), 0, c->stream, G->wbf + G->off.W_AB[l],
                       H2, H, tab);
    hipLaunchKernelGGL(k_transpose, dim3(nb((int64_t)H * H2, 256)), dim3(256), 0, c->stream, G->wbf + G->off.W_U[l],
                       H, H2, tab + (int64_t)H * H2);
  }
  XDEM_HIP(c, hipGetLastError());
  return XDEM_OK;
}

template <int EPI, int M, int K1, int K2>
int gemm_p(xdem_gnn* G, const GemmArgs& a) {
  xdem_ctx* c = G->c;
  if (a.n <= 0) return XDEM_OK;
  constexpr int K = K1 + K2;
  static_assert(M % 128 == 0, "column groups of 128");
  const int lds = (M * (K + 8) + (EPI == EPI_UPDATE_DEC ? 32 * (H + 8) : 0)) * (int)sizeof(uint16_t);
  XDEM_HIP(c, hipFuncSetAttribute((const void*)k_gemm_p<EPI, M, K1, K2>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  const int64_t ntiles = (a.n + 255) / 256;
  const int grid = (int)std::min<int64_t>(ntiles, 512);
  hipLaunchKernelGGL((k_gemm_p<EPI, M, K1, K2>), dim3(grid), dim3(512), lds, c->stream, a);
  XDEM_HIP(c, hipGetLastError());
  return XDEM_OK;
}

int ensure_partial(xdem_gnn* G, int64_t floats) {
  if (floats <= G->cap_partial) return XDEM_OK;
  if (G->cap_partial > 0) return set_err(G->c, XDEM_ERR_CAPACITY, "gnn: partial-sum buffer too small");
  int rc = galloc(G->c, &G->partial, floats);
  if (rc) return rc;
  G->cap_partial = floats;
  return XDEM_OK;
}

// weight gradient of an mg x K matrix into grads[off]
int wgrad(xdem_gnn* G, const uint16_t* gm, int mg, const uint16_t* x1, int k1, const uint16_t* x2, int k2,
          int64_t goff) {
  xdem_ctx* c = G->c;
  const int64_t n = c->n;
  const int K = k1 + k2;
  int blocks = (int)std::min<int64_t>(256, std::max<int64_t>(1, (n + 63) / 64));
  int64_t rpb = (n + blocks - 1) / blocks;
  rpb = (rpb + 15) / 16 * 16;
  blocks = (int)((n + rpb - 1) / rpb);
  const int64_t len = (int64_t)mg * K;
  int rc = ensure_partial(G, len * blocks);
  if (rc) return rc;
  WgradArgs a{n, gm, mg, x1, k1, x2, k2, rpb, G->partial};
  hipLaunchKernelGGL(k_wgrad, dim3(blocks), dim3(256), 0, c->stream, a);
  hipLaunchKernelGGL(k_sum_partials, dim3(nb(len, 256)), dim3(256), 0, c->stream, G->partial, blocks, len,
                     G->g + goff);
  XDEM_HIP(c, hipGetLastError());
  return XDEM_OK;
}
int colsum(xdem_gnn* G, const uint16_t* gm, int m, int64_t goff) {
  xdem_ctx* c = G->c;
  const int blocks = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, c->n));
  int rc = ensure_partial(G, (int64_t)blocks * m);
  if (rc) return rc;
  hipLaunchKernelGGL(k_colsum, dim3(blocks), dim3(256), 0, c->stream, gm, c->n, m, G->partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(nb(m, 256)), dim3(256), 0, c->stream, G->partial, blocks, (int64_t)m,
                     G->g + goff);
  XDEM_HIP(c, hipGetLastError());
  return XDEM_OK;
}

int check_ready(xdem_gnn* G) {
  xdem_ctx* c = G->c;
  if (c->nshapes > 0) return set_err(c, XDEM_ERR_UNSUPPORTED, "surrogate: spheres only");
  if (!c->nl_valid) return set_err(c, XDEM_ERR_STATE, "surrogate: build the neighbour list first");
  if (c->n > 0 && c->n > (int64_t)INT32_MAX / H2) return set_err(c, XDEM_ERR_CAPACITY, "surrogate: too many particles");
  return XDEM_OK;
}

// inputs + forward at the current state
int forward(xdem_gnn* G) {
  xdem_ctx* c = G->c;
  int rc;
  if ((rc = check_ready(G))) return rc;
  if ((rc = engine_flush_kick(c))) return rc;
  if ((rc = ensure_buffers(G))) return rc;
  const int64_t n = c->n;
  G->fwd_valid = false;
  if (n == 0) return XDEM_OK;
  const NbrList& nl = c->nl[c->nl_cur];
  const bool tr = G->cfg.train != 0;
  FeatArgs fa{n, live(c), c->grid, nl, (float)G->cfg.r0, (float)G->cfg.v0, (float)G->cfg.ov_scale, G->x, G->e};
  hipLaunchKernelGGL(k_features, dim3(nb(n, 256)), dim3(256), 0, c->stream, fa);
  if (tr) hipLaunchKernelGGL(k_mirror_all, dim3(nb(n, 256)), dim3(256), 0, c->stream, n, nl, G->mirror);
  EncArgs ea{n, G->x, G->wbf + G->off.W_in, G->wbf + G->off.b_in, G->h[0]};
  hipLaunchKernelGGL(k_encode_mfma, dim3(nb(n, 128)), dim3(256), 0, c->stream, ea);
  XDEM_HIP(c, hipGetLastError());
  for (int l = 0; l < L; l++) {
    uint16_t* hin = tr ? G->h[l] : G->h[l & 1];
    uint16_t* hout = tr ? G->h[l + 1] : G->h[(l + 1) & 1];
    uint16_t* pq = tr ? G->pq[l] : G->pq[0];
    uint16_t* mm = tr ? G->m[l] : G->m[0];
    GemmArgs g1{};
    g1.n = n; g1.a1 = hin; g1.k1 = H; g1.a2 = nullptr; g1.k2 = 0;
    g1.b = G->wbf + G->off.W_AB[l]; g1.m = H2; g1.ybf = pq;
    if ((rc = gemm_p<EPI_BF16, H2, H, 0>(G, g1))) return rc;
    MsgArgs ma{};
    ma.n = n; ma.nl = nl; ma.e = G->e; ma.pq = pq;
    ma.wE = G->wr + G->off.W_E[l]; ma.bE = G->wr + G->off.b_E[l]; ma.m = mm;
    hipLaunchKernelGGL(k_msg_fwd_mfma, dim3((unsigned)((n + 31) / 32)), dim3(256), 0, c->stream, ma,
                       G->wbf + G->off.W_E[l]);
    GemmArgs g2{};
    g2.n = n; g2.a1 = hin; g2.k1 = H; g2.a2 = mm; g2.k2 = H;
    g2.b = G->wbf + G->off.W_U[l]; g2.m = H; g2.bias = G->wr + G->off.b_U[l];
    g2.u_out = tr ? G->u[l] : nullptr; g2.h_out = hout;
    if (!tr && l == L - 1) {
      // inference: the decoder runs on the last hidden state in the update's registers
      g2.h_out = nullptr;
      g2.dec_w = G->wbf + G->off.W_out; g2.dec_b = G->wr + G->off.b_out; g2.dec_out = G->out;
      if ((rc = gemm_p<EPI_UPDATE_DEC, H, H, H>(G, g2))) return rc;
    } else if ((rc = gemm_p<EPI_UPDATE, H, H, H>(G, g2))) {
      return rc;
    }
  }
  if (tr) {
    DecArgs da{n, G->h[L], G->wr + G->off.W_out, G->wbf + G->off.W_out, G->wr + G->off.b_out, G->out, nullptr, nullptr,
               nullptr};
    hipLaunchKernelGGL(k_decode_mfma, dim3(nb(n, 128)), dim3(256), 0, c->stream, da);
  }
  XDEM_HIP(c, hipGetLastError());
  G->fwd_valid = true;
  G->fwd_n = n;
  return XDEM_OK;
}

int backward(xdem_gnn* G, double* loss) {
  xdem_ctx* c = G->c;
  if (!G->cfg.train) return set_err(c, XDEM_ERR_STATE, "gnn: created without training buffers");
  if (!G->fwd_valid || G->fwd_n != c->n) return set_err(c, XDEM_ERR_STATE, "gnn: run the forward pass first");
  const int64_t n = c->n;
  int rc;
  const NbrList& nl = c->nl[c->nl_cur];
  const int wgrid = (int)std::min<int64_t>(nb(n * 64, 256), kRedBlocks);
  // decoder: loss, dh_L, dW_out, db_out
  const int dlen = FO * H + FO + 1;
  if ((rc = ensure_partial(G, (int64_t)wgrid * dlen))) return rc;
  DecArgs da{n, G->h[L], G->wr + G->off.W_out, G->wbf + G->off.W_out, G->wr + G->off.b_out, G->out, G->target, G->dh, G->partial};
  hipLaunchKernelGGL(k_decode<true>, dim3(wgrid), dim3(256), 0, c->stream, da);
  // reduce into g[W_out..b_out] (contiguous) and the loss (one float after)
  hipLaunchKernelGGL(k_sum_partials, dim3(nb(dlen, 256)), dim3(256), 0, c->stream, G->partial, wgrid, (int64_t)dlen,
                     G->dec_sum);
  XDEM_HIP(c, hipMemcpyAsync(G->g + G->off.W_out, G->dec_sum, sizeof(float) * (FO * H + FO), hipMemcpyDeviceToDevice,
                             c->stream));
  XDEM_HIP(c, hipMemcpyAsync(G->h_loss, G->dec_sum + FO * H + FO, sizeof(float), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipGetLastError());
  for (int l = L - 1; l >= 0; l--) {
    // update: du, dW_U, db_U, dX -> dh (residual already in dh), dm
    hipLaunchKernelGGL(k_relu_grad, dim3(nb(n * H, 256)), dim3(256), 0, c->stream, n * H, G->dh, G->u[l], G->du);
    if ((rc = wgrad(G, G->du, H, G->h[l], H, G->m[l], H, G->off.W_U[l]))) return rc;
    if ((rc = colsum(G, G->du, H, G->off.b_U[l]))) return rc;
    uint16_t* tab = G->wT + (int64_t)l * 2 * H * H2;
    GemmArgs g1{};
    g1.n = n; g1.a1 = G->du; g1.k1 = H; g1.k2 = 0;
    g1.b = tab + (int64_t)H * H2;  // W_U^T [2H][H]
    g1.m = H2; g1.dh = G->dh; g1.dm = G->dm;
    if ((rc = gemm_p<EPI_BWD_U, H2, H, 0>(G, g1))) return rc;
    // messages: dP|dQ, dW_E, db_E
    const int mb = wgrid;
    if ((rc = ensure_partial(G, (int64_t)mb * H * (FE + 1)))) return rc;
    MsgArgs ma{};
    ma.n = n; ma.nl = nl; ma.e = G->e; ma.pq = G->pq[l];
    ma.wE = G->wr + G->off.W_E[l]; ma.bE = G->wr + G->off.b_E[l];
    ma.dm = G->dm; ma.mirror = G->mirror; ma.dpq = G->dpq;
    ma.part_w = G->partial; ma.part_b = G->partial + (int64_t)mb * H * FE;
    hipLaunchKernelGGL(k_msg_bwd, dim3(mb), dim3(256), 0, c->stream, ma);
    hipLaunchKernelGGL(k_sum_partials, dim3(nb((int64_t)H * FE, 256)), dim3(256), 0, c->stream, ma.part_w, mb,
                       (int64_t)H * FE, G->g + G->off.W_E[l]);
    hipLaunchKernelGGL(k_sum_partials, dim3(nb(H, 256)), dim3(256), 0, c->stream, ma.part_b, mb, (int64_t)H,
                       G->g + G->off.b_E[l]);
    XDEM_HIP(c, hipGetLastError());
    // P|Q: dW_AB, dh += dPQ W_AB
    if ((rc = wgrad(G, G->dpq, H2, G->h[l], H, nullptr, 0, G->off.W_AB[l]))) return rc;
    GemmArgs g2{};
    g2.n = n; g2.a1 = G->dpq; g2.k1 = H2; g2.k2 = 0;
    g2.b = tab;  // W_AB^T [H][2H]
    g2.m = H; g2.dh = G->dh;
    if ((rc = gemm_p<EPI_ADD, H, H2, 0>(G, g2))) return rc;
  }
  // encoder
  {
    const int eb = wgrid;
    const int64_t len = (int64_t)H * (FN + 1);
    if ((rc = ensure_partial(G, eb * len + len))) return rc;
    hipLaunchKernelGGL(k_encode_bwd, dim3(eb), dim3(256), 0, c->stream, n, G->x, G->h[0], G->dh, G->partial);
    float* sum = G->partial + eb * len;
    hipLaunchKernelGGL(k_sum_partials, dim3(nb(len, 256)), dim3(256), 0, c->stream, G->partial, eb, len, sum);
    hipLaunchKernelGGL(k_split_wb, dim3(nb(len, 256)), dim3(256), 0, c->stream, sum, H, FN, G->g + G->off.W_in,
                       G->g + G->off.b_in);
    XDEM_HIP(c, hipGetLastError());
  }
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  if (loss) *loss = (double)G->h_loss[0] / (double)(n * FO);
  return XDEM_OK;
}

}  // namespace

// ------------------------------------------------------------------------------------------ C ABI
extern "C" int xdem_gnn_create(xdem_ctx* c, const xdem_gnn_config* cfg, xdem_gnn** out) {
  if (!c || !cfg || !out) return XDEM_ERR_ARG;
  *out = nullptr;
  if (!(cfg->r0 > 0) || !(cfg->v0 > 0) || !(cfg->a0 > 0) || !(cfg->m0 > 0) || cfg->substeps < 1)
    return set_err(c, XDEM_ERR_ARG, "gnn: scales must be > 0 and substeps >= 1");
  XDEM_HIP(c, hipSetDevice(c->device));
  xdem_gnn* G = new xdem_gnn();
  G->c = c;
  G->cfg = *cfg;
  G->off = offsets();
  const int64_t np = G->off.total;
  int rc;
  if ((rc = galloc(c, &G->p, np)) || (rc = galloc(c, &G->g, np)) || (rc = galloc(c, &G->m1, np)) ||
      (rc = galloc(c, &G->m2, np)) || (rc = galloc(c, &G->wbf, np)) || (rc = galloc(c, &G->wr, np)) ||
      (rc = galloc(c, &G->wT, (int64_t)L * 2 * H * H2)) || (rc = galloc(c, &G->dpart, 4 * 65536)) ||
      (rc = galloc(c, &G->dec_sum, FO * H + FO + 1)) || (rc = ensure_partial(G, kPartialFloats))) {
    delete G;
    return rc;
  }
  XDEM_HIP(c, hipMemsetAsync(G->p, 0, sizeof(float) * np, c->stream));
  XDEM_HIP(c, hipMemsetAsync(G->g, 0, sizeof(float) * np, c->stream));
  XDEM_HIP(c, hipMemsetAsync(G->m1, 0, sizeof(float) * np, c->stream));
  XDEM_HIP(c, hipMemsetAsync(G->m2, 0, sizeof(float) * np, c->stream));
  hipLaunchKernelGGL(k_round_params, dim3(nb(np, 256)), dim3(256), 0, c->stream, np, G->p, G->wbf, G->wr);
  if ((rc = refresh_weights(G))) return rc;
  XDEM_HIP(c, hipHostMalloc((void**)&G->h_dpart, sizeof(double) * 4 * 65536, 0));
  XDEM_HIP(c, hipHostMalloc((void**)&G->h_loss, sizeof(float) * 4, 0));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  *out = G;
  return XDEM_OK;
}

extern "C" int xdem_gnn_destroy(xdem_gnn* G) {
  if (!G) return XDEM_OK;
  (void)hipSetDevice(G->c->device);
  (void)hipStreamSynchronize(G->c->stream);
  void* bufs[] = {G->p,  G->g,   G->m1,     G->m2,      G->wbf,     G->wr,      G->wT,
                  G->x,  G->e,   G->out,    G->target,  G->dh,      G->dm,      G->du,
                  G->dpq, G->mirror, G->partial, G->snap, G->xs, G->xf, G->dpart, G->dec_sum};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (int k = 0; k <= L; k++)
    if (G->h[k]) (void)hipFree(G->h[k]);
  for (int k = 0; k < L; k++) {
    if (G->pq[k]) (void)hipFree(G->pq[k]);
    if (G->m[k]) (void)hipFree(G->m[k]);
    if (G->u[k]) (void)hipFree(G->u[k]);
  }
  if (G->h_dpart) (void)hipHostFree(G->h_dpart);
  if (G->h_loss) (void)hipHostFree(G->h_loss);
  delete G;
  return XDEM_OK;
}

extern "C" int xdem_gnn_set_params(xdem_gnn* G, const float* params, int64_t n) {
  if (!G || !params) return XDEM_ERR_ARG;
  xdem_ctx* c = G->c;
  if (n != G->off.total) return set_err(c, XDEM_ERR_ARG, "gnn: parameter count must be XDEM_GNN_NPARAMS");
  XDEM_HIP(c, hipMemcpyAsync(G->p, params, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  XDEM_HIP(c, hipMemsetAsync(G->m1, 0, sizeof(float) * n, c->stream));
  XDEM_HIP(c, hipMemsetAsync(G->m2, 0, sizeof(float) * n, c->stream));
  G->adam_t = 0;
  hipLaunchKernelGGL(k_round_params, dim3(nb(n, 256)), dim3(256), 0, c->stream, n, G->p, G->wbf, G->wr);
  int rc = refresh_weights(G);
  if (rc) return rc;
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  return XDEM_OK;
}

extern "C" int xdem_gnn_get_params(xdem_gnn* G, float* params, int64_t n) {
  if (!G || !params || n != G->off.total) return XDEM_ERR_ARG;
  XDEM_HIP(G->c, hipMemcpyAsync(params, G->p, sizeof(float) * n, hipMemcpyDeviceToHost, G->c->stream));
  XDEM_HIP(G->c, hipStreamSynchronize(G->c->stream));
  return XDEM_OK;
}

extern "C" int xdem_gnn_get_grads(xdem_gnn* G, float* grads, int64_t n) {
  if (!G || !grads || n != G->off.total) return XDEM_ERR_ARG;
  XDEM_HIP(G->c, hipMemcpyAsync(grads, G->g, sizeof(float) * n, hipMemcpyDeviceToHost, G->c->stream));
  XDEM_HIP(G->c, hipStreamSynchronize(G->c->stream));
  return XDEM_OK;
}

extern "C" int xdem_gnn_forward(xdem_gnn* G, float* out) {
  if (!G) return XDEM_ERR_ARG;
  int rc = forward(G);
  if (rc) return rc;
  xdem_ctx* c = G->c;
  if (out && c->n > 0)
    XDEM_HIP(c, hipMemcpyAsync(out, G->out, sizeof(float) * FO * c->n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  return XDEM_OK;
}

extern "C" int xdem_gnn_set_target(xdem_gnn* G, const float* target) {
  if (!G || !target) return XDEM_ERR_ARG;
  xdem_ctx* c = G->c;
  if (!G->cfg.train) return set_err(c, XDEM_ERR_STATE, "gnn: created without training buffers");
  int rc = ensure_buffers(G);
  if (rc) return rc;
  if (c->n > 0)
    XDEM_HIP(c, hipMemcpyAsync(G->target, target, sizeof(float) * FO * c->n, hipMemcpyHostToDevice, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  return XDEM_OK;
}

extern "C" int xdem_gnn_backward(xdem_gnn* G, double* loss) {
  if (!G) return XDEM_ERR_ARG;
  if (G->c->n == 0) {
    if (loss) *loss = 0;
    return XDEM_OK;
  }
  return backward(G, loss);
}

extern "C" int xdem_gnn_adam(xdem_gnn* G, double lr, double beta1, double beta2, double eps) {
  if (!G) return XDEM_ERR_ARG;
  xdem_ctx* c = G->c;
  G->adam_t++;
  const double c1 = 1.0 - std::pow(beta1, (double)G->adam_t), c2 = 1.0 - std::pow(beta2, (double)G->adam_t);
  const int64_t np = G->off.total;
  hipLaunchKernelGGL(k_adam, dim3(nb(np, 256)), dim3(256), 0, c->stream, np, G->p, G->g, G->m1, G->m2, (float)lr,
                     (float)beta1, (float)beta2, (float)eps, (float)c1, (float)c2, G->wbf, G->wr);
  XDEM_HIP(c, hipGetLastError());
  int rc = refresh_weights(G);
  if (rc) return rc;
  G->fwd_valid = false;
  return XDEM_OK;
}

extern "C" int xdem_gnn_train_window(xdem_gnn* G, const xdem_scheme* s, double lr, double* loss, int32_t* used) {
  if (!G || !s) return XDEM_ERR_ARG;
  xdem_ctx* c = G->c;
  if (!G->cfg.train) return set_err(c, XDEM_ERR_STATE, "gnn: created without training buffers");
  if (used) *used = 0;
  if (loss) *loss = 0;
  int rc;
  if ((rc = forward(G))) return rc;
  const int64_t n = c->n;
  if (n == 0) return XDEM_OK;
  hipLaunchKernelGGL(k_snapshot, dim3(nb(n, 256)), dim3(256), 0, c->stream, n, live(c), G->snap);
  xdem_run_stats st{};
  if ((rc = xdem_run_steps(c, s, G->cfg.substeps, &st))) return rc;
  if (st.rebuilds > 0) {
    G->fwd_valid = false;
    return XDEM_OK;
  }
  const V3 g = s->gravity_on ? V3{s->gravity[0], s->gravity[1], s->gravity[2]} : V3{0, 0, 0};
  hipLaunchKernelGGL(k_targets, dim3(nb(n, 256)), dim3(256), 0, c->stream, n, live(c), G->snap,
                     s->dt * G->cfg.substeps, g, force_scale(G), torque_scale(G), G->target);
  XDEM_HIP(c, hipGetLastError());
  if ((rc = backward(G, loss))) return rc;
  if ((rc = xdem_gnn_adam(G, lr, 0.9, 0.999, 1e-8))) return rc;
  if (used) *used = 1;
  return XDEM_OK;
}

extern "C" int xdem_gnn_dump_graph(xdem_gnn* G, int64_t cap, uint32_t* src, uint32_t* dst, uint16_t* edge_in,
                                   uint16_t* node_in, int64_t* n_edges) {
  if (!G || !src || !dst || !edge_in || !node_in) return XDEM_ERR_ARG;
  xdem_ctx* c = G->c;
  if (n_edges) *n_edges = 0;
  if (!G->fwd_valid) return set_err(c, XDEM_ERR_STATE, "gnn: run the forward pass first");
  const int64_t n = c->n;
  const NbrList& L_ = c->nl[c->nl_cur];
  const int64_t nw = (n + kWave - 1) / kWave;
  std::vector<uint32_t> cnt(n), word(L_.slots);
  std::vector<uint64_t> base(nw + 1);
  std::vector<uint16_t> e(L_.slots * FE);
  XDEM_HIP(c, hipMemcpyAsync(cnt.data(), L_.cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(base.data(), L_.base, sizeof(uint64_t) * (nw + 1), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(word.data(), L_.word, sizeof(uint32_t) * L_.slots, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(e.data(), G->e, sizeof(uint16_t) * FE * L_.slots, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipMemcpyAsync(node_in, G->x, sizeof(uint16_t) * FN * n, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  int64_t k = 0;
  for (int64_t x = 0; x < n; x++) {
    const uint64_t b = base[x >> 6] + (uint64_t)(x & 63);
    for (uint32_t q = 0; q < cnt[x]; q++) {
      const uint64_t sl = b + (uint64_t)q * kWave;
      if (k >= cap) return set_err(c, XDEM_ERR_CAPACITY, "gnn dump: cap too small");
      src[k] = (uint32_t)x;
      dst[k] = word[sl] & kIdxMask;
      memcpy(edge_in + k * FE, e.data() + sl * FE, sizeof(uint16_t) * FE);
      k++;
    }
  }
  if (n_edges) *n_edges = k;
  return XDEM_OK;
}

namespace {
// displacement test against the rebuild backup; rebuilds when over
int maybe_rebuild(xdem_gnn* G, const xdem_scheme* s, int64_t* rebuilds) {
  xdem_ctx* c = G->c;
  int rc;
  const double thr = 0.5 * s->rcut_inc;
  XDEM_HIP(c, hipMemsetAsync(c->d_flags, 0, sizeof(int32_t), c->stream));
  if ((rc = launch_displ(c, thr * thr))) return rc;
  XDEM_HIP(c, hipMemcpyAsync(c->h_flags, c->d_flags, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  if (c->h_flags[0]) {
    XDEM_HIP(c, hipMemsetAsync(c->d_flags, 0, 4 * sizeof(int32_t), c->stream));
    if ((rc = engine_rebuild(c, s->rcut_inc))) return rc;
    (*rebuilds)++;
  }
  return XDEM_OK;
}

int drift(xdem_gnn* G, double* e_sur, double* e_free, double* r_sur, double* r_free) {
  xdem_ctx* c = G->c;
  const int64_t n = c->n;
  const int blocks = nb(n, 256);
  if (blocks > 65536) return set_err(c, XDEM_ERR_CAPACITY, "surrogate: too many particles for the drift check");
  hipLaunchKernelGGL(k_drift, dim3(blocks), dim3(256), 0, c->stream, n, live(c), G->xs, G->xf, G->dpart);
  XDEM_HIP(c, hipMemcpyAsync(G->h_dpart, G->dpart, sizeof(double) * 4 * blocks, hipMemcpyDeviceToHost, c->stream));
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  double a = 0, b = 0, sa = 0, sb = 0;
  for (int k = 0; k < blocks; k++) {
    a = std::max(a, G->h_dpart[4 * k]);
    b = std::max(b, G->h_dpart[4 * k + 1]);
    sa += G->h_dpart[4 * k + 2];
    sb += G->h_dpart[4 * k + 3];
  }
  *e_sur = a;
  *e_free = b;
  *r_sur = n > 0 ? std::sqrt(sa / (double)n) : 0.0;
  *r_free = n > 0 ? std::sqrt(sb / (double)n) : 0.0;
  return XDEM_OK;
}
}  // namespace

extern "C" int xdem_surrogate_run(xdem_gnn* G, const xdem_scheme* s, int64_t nsteps, int32_t check_every, double tol,
                                  xdem_surrogate_stats* stats) {
  if (!G || !s || nsteps < 0 || check_every < 1) return XDEM_ERR_ARG;
  xdem_ctx* c = G->c;
  if (stats) *stats = {};
  if (!c->ids_dense) return set_err(c, XDEM_ERR_UNSUPPORTED, "surrogate: ids must be a permutation of 0..n-1");
  if (s->form != XDEM_FORM_FP64) return set_err(c, XDEM_ERR_UNSUPPORTED, "surrogate: DEM checks run in FP64");
  int rc;
  if ((rc = check_ready(G))) return rc;
  xdem_surrogate_stats st{};
  const int K = G->cfg.substeps;
  const double T = s->dt * K;
  const V3 g = s->gravity_on ? V3{s->gravity[0], s->gravity[1], s->gravity[2]} : V3{0, 0, 0};
  int64_t dem_until = -1;  // steps below this index run plain DEM (after a failed check)
  bool force_stale = false;
  for (int64_t k = 0; k < nsteps; k++) {
    const bool check = (k % check_every) == check_every - 1;
    if (k < dem_until || check) {
      // DEM for this step (checked or fallback); forces first if surrogate steps moved the state
      if (check) {
        if ((rc = forward(G))) return rc;
        hipLaunchKernelGGL(k_sur_advance, dim3(nb(c->n, 256)), dim3(256), 0, c->stream, c->n, live(c), G->out, T, g,
                           force_scale(G), torque_scale(G), 0, 1, G->xs, G->xf);
        XDEM_HIP(c, hipGetLastError());
      }
      if (force_stale) {
        if ((rc = engine_force_pass(c, s))) return rc;
        c->pending_kick = false;
        force_stale = false;
      }
      xdem_run_stats rs{};
      if ((rc = xdem_run_steps(c, s, K, &rs))) return rc;
      st.dem_steps += K;
      st.rebuilds += rs.rebuilds;
      if (check) {
        double e1 = 0, e2 = 0, q1 = 0, q2 = 0;
        if ((rc = drift(G, &e1, &e2, &q1, &q2))) return rc;
        st.checks++;
        st.max_err = std::max(st.max_err, e1);
        st.max_err_free = std::max(st.max_err_free, e2);
        st.rms_err = std::max(st.rms_err, q1);
        st.rms_err_free = std::max(st.rms_err_free, q2);
        if (e1 > tol) {
          st.fallbacks++;
          dem_until = k + 1 + check_every;
        }
      }
      continue;
    }
    if ((rc = forward(G))) return rc;
    hipLaunchKernelGGL(k_sur_advance, dim3(nb(c->n, 256)), dim3(256), 0, c->stream, c->n, live(c), G->out, T, g,
                       force_scale(G), torque_scale(G), 1, 0, G->xs, G->xf);
    XDEM_HIP(c, hipGetLastError());
    c->cells_valid = false;
    force_stale = true;
    st.surrogate_steps++;
    if ((rc = maybe_rebuild(G, s, &st.rebuilds))) return rc;
  }
  if (force_stale) {
    if ((rc = engine_force_pass(c, s))) return rc;
    c->pending_kick = false;
  }
  XDEM_HIP(c, hipStreamSynchronize(c->stream));
  if (stats) *stats = st;
  return XDEM_OK;
}
