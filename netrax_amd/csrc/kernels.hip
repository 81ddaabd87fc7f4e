// HIP kernels of the network-likelihood engine (gfx950).
//
// Data layout in HBM (one engine):
//   CLV arena   real [n_variants][n_sites_pad][C][S]  site-major, state innermost
//   scalers     int32 [n_variants][n_sites_pad]
//   adjoints    real  [n_variants][n_sites_pad][C][S]
//   tips        uint8 (S <= 8) or uint32 (S = 20) state masks [n_tips][n_sites_pad]
//   P, P', P''  real  [n_parts][n_edges][C][S][S]
// Every site-parallel kernel works on tiles of TILE = 64 sites; a workgroup
// holds TILE * C threads, thread = (site, rate category), so the C lanes of
// a site are adjacent and a CLV row of a workgroup is one contiguous
// TILE*C*S*sizeof(real) span.  Workgroup b of a launch over (tile, op)
// handles tile b % n_tiles_grid of op b / n_tiles_grid; n_tiles_grid is a
// multiple of 8, so a tile is always dispatched to the same XCD (round-robin
// over workgroup ids) and its children's CLVs stay in that XCD's L2.
#include <cfloat>
#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>
#include <type_traits>

#include "engine_state.hpp"

namespace nrx {

void HipCheck::check(hipError_t err, const char *what) {
  if (err != hipSuccess)
    throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(err) + " in " + what);
}

namespace {

// libpll's per-site scaling constants (PLL_SCALE_THRESHOLD = 2^-256,
// PLL_SCALE_FACTOR = 2^256): a site is rescaled when all its entries are
// below the threshold; the scaler counts rescalings.
constexpr double SCALE_THRESHOLD = 0x1p-256;
constexpr double SCALE_FACTOR = 0x1p256;
constexpr double LOG_SCALE_THRESHOLD = -256.0 * 0.69314718055994530941723212145818;

template <int S>
struct TipType {
  using type = uint8_t;
};
template <>
struct TipType<20> {
  using type = uint32_t;
};

// Dynamic LDS is one untyped buffer per kernel (all extern __shared__
// declarations of a translation unit must agree on the type).
template <typename T>
__device__ __forceinline__ T *dyn_lds() {
  extern __shared__ double lds_dyn[];
  return reinterpret_cast<T *>(lds_dyn);
}

// Pointers read from a device-side descriptor table are generic (flat)
// pointers to the compiler; kernels that walk such tables retype them as
// global-memory pointers so CLV traffic uses global (not flat) loads and
// stores with scalar base addresses.
// (The host pass of this file never runs these kernels and has no address
// spaces, so the qualifier exists in the device pass only.)
#ifdef __HIP_DEVICE_COMPILE__
#define GAS __attribute__((address_space(1)))
#define CAS __attribute__((address_space(4)))
#else
#define GAS
#define CAS
#endif
template <typename T>
__device__ __forceinline__ GAS T *to_global(T *p) {
  return (GAS T *)p;
}
// Op tables and group bounds never change during a launch: read through
// the constant address space they are fetched with scalar loads, whatever
// stores the kernel makes in between.
template <typename T>
__device__ __forceinline__ const CAS T *to_constant(const T *p) {
  return (const CAS T *)p;
}

// A thread's word in one slot of a per-slot arena (CLVs, adjoints,
// scalers, tip rows): the slot's base address is the same for the whole
// wave (scalar arithmetic) and the thread's offset into every slot is one
// fixed 32-bit byte offset, so each access is a scalar base plus that
// offset, with no per-lane 64-bit address arithmetic.
template <typename T>
__device__ __forceinline__ T *at(T *slot_base, uint32_t byte_off) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T *>(reinterpret_cast<B *>(slot_base) + byte_off);
}
#ifdef __HIP_DEVICE_COMPILE__
template <typename T>
__device__ __forceinline__ GAS T *at(GAS T *slot_base, uint32_t byte_off) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<GAS T *>(reinterpret_cast<GAS B *>(slot_base) + byte_off);
}
#endif

// Cross-lane moves by DPP (a VALU operand modifier, no LDS round trip).
// Controls: quad_perm(p0, p1, p2, p3) = p0 | p1 << 2 | p2 << 4 | p3 << 6
// within each group of 4 lanes; row_shr:n = 0x110 + n shifts up by n lanes
// within each row of 16.  Lanes whose source lies outside the row read 0.
constexpr int DPP_XOR1 = 0xB1;  // quad_perm(1, 0, 3, 2): lane ^ 1
constexpr int DPP_XOR2 = 0x4E;  // quad_perm(2, 3, 0, 1): lane ^ 2
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// value of lane `off` apart within the C adjacent lanes of a site
template <int OFF, typename T>
__device__ __forceinline__ T lane_xor(T v) {
  if constexpr (OFF == 1)
    return dpp<DPP_XOR1>(v);
  else if constexpr (OFF == 2)
    return dpp<DPP_XOR2>(v);
  else
    return __shfl_xor(v, OFF);
}

template <int C>
__device__ __forceinline__ double cat_max(double m) {
  if constexpr (C > 1) m = fmax(m, lane_xor<1>(m));
  if constexpr (C > 2) m = fmax(m, lane_xor<2>(m));
  if constexpr (C > 4) m = fmax(m, lane_xor<4>(m));
  return m;
}
template <int C>
__device__ __forceinline__ double cat_sum(double m) {
  if constexpr (C > 1) m += lane_xor<1>(m);
  if constexpr (C > 2) m += lane_xor<2>(m);
  if constexpr (C > 4) m += lane_xor<4>(m);
  return m;
}

// Waits until every vector-memory load this wave issued has returned
// (s_waitcnt vmcnt(0); expcnt and lgkmcnt left unconstrained).  Placed after
// a load phase whose loads depend on runtime conditions, it leaves every
// path with no pending load, so the compiler needs no conservative waits
// before the next phase's loads.
__device__ __forceinline__ void wait_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }
// Load phases are bracketed by s_setprio 1 / s_setprio 0.  Measured 1-1.5 %
// faster in the batched and plain passes; swapping the two priorities
// measured the same, so the gain comes from the fixed instruction order the
// pair imposes (the compiler moves nothing across an s_setprio), not from
// the wave arbitration itself.

__device__ __forceinline__ double readlane(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over the 64 lanes of a wave, in a fixed order (deterministic); the
// result is the same in every lane.  Row sums by four row_shr steps (lane
// 15 of each row of 16 ends with its row's sum S_r), then row_bcast:15
// (0x142: each row adds lane 15 of the row before it; lane 31 = S0 + S1,
// lane 63 = S2 + S3) and row_bcast:31 (0x143: rows 2 and 3 add lane 31), so
// lane 63 ends with the total: one lane read instead of four.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp<0x111>(v);
  v += dpp<0x112>(v);
  v += dpp<0x114>(v);
  v += dpp<0x118>(v);
  v += dpp<0x142>(v);
  v += dpp<0x143>(v);
  return readlane(v, 63);
}

// Deterministic workgroup sum; result valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) r += sh[i];
  }
  __syncthreads();
  return r;
}

// Gradient partials: one value per (site block, gradient row, wave of the
// block), stored [block][row][wave].  Rows are numbered in the order the
// adjoint walks the ops, so a workgroup's waves fill whole cache lines of
// their own block instead of scattering 8-byte writes over rows shared with
// other workgroups (partial-line writes cost a read-modify-write in HBM).
__host__ __device__ __forceinline__ size_t grad_index(int64_t block, int n_uses, int g, int pw, int w) {
  return (static_cast<size_t>(block) * n_uses + g) * pw + w;
}

template <int S>
__device__ __forceinline__ void load_clv(const double *__restrict__ p, double (&x)[S]) {
  if constexpr (S == 4) {
    const double2 a = reinterpret_cast<const double2 *>(p)[0];
    const double2 b = reinterpret_cast<const double2 *>(p)[1];
    x[0] = a.x;
    x[1] = a.y;
    x[2] = b.x;
    x[3] = b.y;
  } else {
#pragma unroll
    for (int i = 0; i < S; ++i) x[i] = p[i];
  }
}

template <int S>
__device__ __forceinline__ void store_clv(double *__restrict__ p, const double (&x)[S]) {
  if constexpr (S == 4) {
    reinterpret_cast<double2 *>(p)[0] = make_double2(x[0], x[1]);
    reinterpret_cast<double2 *>(p)[1] = make_double2(x[2], x[3]);
  } else {
#pragma unroll
    for (int i = 0; i < S; ++i) p[i] = x[i];
  }
}

template <int S>
__device__ __forceinline__ void tip_vec(uint32_t mask, double (&x)[S]) {
#pragma unroll
  for (int i = 0; i < S; ++i) x[i] = ((mask >> i) & 1u) ? 1.0 : 0.0;
}

// y = M x for one S x S row-major matrix
template <int S>
__device__ __forceinline__ void matvec(const double *__restrict__ M, const double (&x)[S],
                                       double (&y)[S]) {
#pragma unroll
  for (int i = 0; i < S; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < S; ++j) s = fma(M[i * S + j], x[j], s);
    y[i] = s;
  }
}

// y = M^T x
template <int S>
__device__ __forceinline__ void matvec_t(const double *__restrict__ M, const double (&x)[S],
                                         double (&y)[S]) {
#pragma unroll
  for (int j = 0; j < S; ++j) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < S; ++i) s = fma(M[i * S + j], x[i], s);
    y[j] = s;
  }
}

// ------------------------------------------------------------- P matrices
// P_{p,e,c} = V diag(exp(lambda r_c t_e)) V^-1, with first and second
// derivatives in t (libpll's pll_update_prob_matrices computes the same
// product from the same eigendecomposition).
__global__ void k_pmatrices(DevLayout L, const double *__restrict__ V,
                            const double *__restrict__ Vinv, const double *__restrict__ ev,
                            const double *__restrict__ rates, const double *__restrict__ brlen,
                            double *__restrict__ P, double *__restrict__ dP,
                            double *__restrict__ d2P, float *__restrict__ Pf,
                            float *__restrict__ dPf) {
  const int S = L.S, C = L.C;
  const int64_t total = static_cast<int64_t>(L.n_parts) * L.n_edges * C * S * S;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int j = idx % S;
  const int i = (idx / S) % S;
  const int c = (idx / (S * S)) % C;
  const int e = (idx / (S * S * C)) % L.n_edges;
  const int p = idx / (static_cast<int64_t>(S) * S * C * L.n_edges);
  const double rt = rates[p * C + c];
  const double t = brlen[e];
  const double *Vp = V + static_cast<size_t>(p) * S * S;
  const double *Wp = Vinv + static_cast<size_t>(p) * S * S;
  const double *lp = ev + static_cast<size_t>(p) * S;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int k = 0; k < S; ++k) {
    const double lr = lp[k] * rt;
    const double x = Vp[i * S + k] * exp(lr * t) * Wp[k * S + j];
    a0 += x;
    a1 += x * lr;
    a2 += x * lr * lr;
  }
  P[idx] = a0;
  dP[idx] = a1;
  d2P[idx] = a2;
  if (Pf) Pf[idx] = static_cast<float>(a0);
  if (dPf) dPf[idx] = static_cast<float>(a1);
}

// ------------------------------------------------------- forward (FP64)
// Plain form: a workgroup per tile walks the levels in one launch
// (k_forward_walk), or, under a memory budget, one launch per level on a
// 2-D grid of (tile, group of ops of the level).  In the latter blockIdx.x
// is the tile and gridDim.x a multiple of 8,
// so every workgroup of a tile lands on the same XCD (workgroup ids are
// dealt round-robin to the 8 XCDs) and finds the children it reads, written
// by the same tile in earlier levels, in that XCD's L2.  For 4-state data
// the P matrices of the group's ops are staged in LDS first: row 2j + k
// holds op j's child-k edge.
template <int S, int C>
__device__ __forceinline__ void stage_ops_p(double *sP, const double *__restrict__ P,
                                            const DevOp *__restrict__ ops, int n_staged, int part,
                                            const DevLayout &L) {
  if constexpr (S == 4) {
    constexpr int ROW = C * S * S;
    for (int i = threadIdx.x; i < n_staged * 2 * ROW; i += blockDim.x) {
      const int slot = i / ROW, r = i % ROW;
      const int j = slot >> 1, k = slot & 1;
      if (k < ops[j].nk)
        sP[i] = P[(static_cast<size_t>(part) * L.n_edges + ops[j].edge[k]) * ROW + r];
    }
    __syncthreads();
  }
}

// P row block of op j's child k for category c: staged in LDS, or read from
// memory (20 states, or ops of an adjoint group beyond the staged ones)
template <int S, int C>
__device__ __forceinline__ const double *op_p(const double *sP, int n_staged,
                                              const double *__restrict__ P, int part,
                                              const DevLayout &L, int j, int k, int edge, int c) {
  if constexpr (S == 4)
    if (j < n_staged) return sP + ((j * 2 + k) * C + c) * (S * S);
  return P + ((static_cast<size_t>(part) * L.n_edges + edge) * C + c) * (S * S);
}

// ops[0, nb) of one pruning level for one tile: all children of the ops are
// loaded before any is combined; their P matrices are staged in sP.
template <int S, int C, int G>
__device__ __forceinline__ void forward_ops(const DevOp *__restrict__ ops, int nb, const DevLayout &L, int part,
                                            int tile, double *__restrict__ clv, int32_t *__restrict__ scl,
                                            const typename TipType<S>::type *__restrict__ tips,
                                            const double *__restrict__ P, double *sP) {
  const int c = threadIdx.x % C;
  const int64_t site = static_cast<int64_t>(tile) * TILE + threadIdx.x / C;
  const int64_t n = L.n_sites_pad;
  {
    // load phase: raw values only (CLV rows, scaler words, tip masks), so no
    // load has to wait for the ones issued before it
    __builtin_amdgcn_s_setprio(1);
    double x[G][2][S];
    int sk[G][2];
    uint32_t tm[G][2];
#pragma unroll
    for (int j = 0; j < G; ++j) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (j < nb) {
          const DevOp op = ops[j];
          const int32_t kid = op.kid[k];
          if (k < op.nk && kid >= 0) {
            load_clv<S>(clv + (static_cast<size_t>(kid) * n + site) * (C * S) + c * S, x[j][k]);
            sk[j][k] = scl[static_cast<size_t>(kid) * n + site];
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (j < nb) {
          const DevOp op = ops[j];
          const int32_t kid = op.kid[k];
          if (k < op.nk && kid < 0) tm[j][k] = tips[static_cast<size_t>(-kid - 1) * n + site];
        }
      }
    }
    // the P matrices after the children's loads: one round trip for both
    __builtin_amdgcn_s_setprio(0);
    stage_ops_p<S, C>(sP, P, ops, nb, part, L);  // ends with a barrier (4 states)
    wait_loads();
#pragma unroll
    for (int j = 0; j < G; ++j) {
      if (j < nb) {
        const DevOp op = ops[j];
        int sc_j = 0;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          if (k < op.nk) {
            if (op.kid[k] < 0)
              tip_vec<S>(tm[j][k], x[j][k]);
            else
              sc_j += sk[j][k];
          }
        }
        double v[S];
#pragma unroll
        for (int i = 0; i < S; ++i) v[i] = 1.0;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          if (k < op.nk) {
            double y[S];
            matvec<S>(op_p<S, C>(sP, nb, P, part, L, j, k, op.edge[k], c), x[j][k], y);
#pragma unroll
            for (int i = 0; i < S; ++i) v[i] *= y[i];
          }
        }
        double m = v[0];
#pragma unroll
        for (int i = 1; i < S; ++i) m = fmax(m, v[i]);
        m = cat_max<C>(m);
        int scj = sc_j;
        if (m < SCALE_THRESHOLD) {
#pragma unroll
          for (int i = 0; i < S; ++i) v[i] *= SCALE_FACTOR;
          scj += 1;
        }
        store_clv<S>(clv + (static_cast<size_t>(op.out) * n + site) * (C * S) + c * S, v);
        if (c == 0) scl[static_cast<size_t>(op.out) * n + site] = scj;
      }
    }
  }
}

// Workgroup (tile, g) computes ops [gG, gG + G) of the launch for one tile.
template <int S, int C, int G>
__global__ void __launch_bounds__(TILE *C)
    k_forward_level(const DevOp *__restrict__ ops, int n_ops, DevLayout L,
                    double *__restrict__ clv, int32_t *__restrict__ scl,
                    const typename TipType<S>::type *__restrict__ tips,
                    const double *__restrict__ P, const int32_t *__restrict__ tile_part) {
  __shared__ double sP[S == 4 ? G * 2 * C * S * S : 1];
  const int tile = blockIdx.x;
  if (tile >= L.n_tiles) return;  // grid padded to a multiple of 8 tiles
  const int g0 = blockIdx.y * G;
  forward_ops<S, C, G>(ops + g0, min(G, n_ops - g0), L, tile_part[tile], tile, clv, scl, tips, P, sP);
}

// The whole forward pass (or its invalid CLVs) in one launch: a workgroup
// per tile walks the levels of `walk` (op ranges of `ops`, lowest level
// first) and each level's ops G at a time.  A CLV and its scaler are read
// back only by the lanes of the wave that stored them (a site's C lanes
// share a wave), so no level boundary needs a launch.
template <int S, int C, int G>
__global__ void __launch_bounds__(TILE *C, 3)
    k_forward_walk(const DevOp *__restrict__ ops, const LevelChunk *__restrict__ walk, int n_walk, DevLayout L,
                   double *__restrict__ clv, int32_t *__restrict__ scl,
                   const typename TipType<S>::type *__restrict__ tips,
                   const double *__restrict__ P, const int32_t *__restrict__ tile_part) {
  __shared__ double sP[S == 4 ? G * 2 * C * S * S : 1];
  const int tile = blockIdx.x;
  if (tile >= L.n_tiles) return;  // grid padded to a multiple of 8 tiles
  const int part = tile_part[tile];
  bool staged = false;
  for (int w = 0; w < n_walk; ++w) {
    const LevelChunk lv = walk[w];
    for (int g0 = lv.op_begin; g0 < lv.op_end; g0 += G) {
      if (S == 4 && staged) __syncthreads();  // every lane is done with the previous group's staged P
      staged = true;
      forward_ops<S, C, G>(ops + g0, min(G, lv.op_end - g0), L, part, tile, clv, scl, tips, P, sP);
    }
  }
}

// ---------------------------------------------------- root site sums (FP64)
// partial[r][wave tile] = sum over the wave's sites of
//   w_s * (log(sum_c (1/C) sum_i pi_i CLV[s,c,i]) + scaler_s * log(2^-256))
// (pll_compute_root_loglikelihood semantics, ImprovedLoglikelihood.cpp:450).
template <int S, int C>
__global__ void __launch_bounds__(TILE *C)
    k_root_partials(const int32_t *__restrict__ root_slots, int n_root, DevLayout L,
                    const double *__restrict__ clv, const int32_t *__restrict__ scl,
                    const double *__restrict__ freqs, const double *__restrict__ weights,
                    const int32_t *__restrict__ tile_part, double *__restrict__ partials,
                    double *__restrict__ seed, const int32_t *__restrict__ root_index) {
  constexpr int PW = TILE * C / 64;
  const int tile = blockIdx.x;
  if (tile >= L.n_tiles) return;
  const int c = threadIdx.x % C;
  const int64_t site = static_cast<int64_t>(tile) * TILE + threadIdx.x / C;
  const int part = tile_part[tile];
  const int64_t n = L.n_sites_pad;
  const size_t n_wt = static_cast<size_t>(L.n_tiles) * PW;
  const int wt = tile * PW + threadIdx.x / 64;
  const double w = weights[site];
  // one root CLV per workgroup (grid tiles x roots): the roots' site sums
  // are independent, and a 2-D grid keeps every CU busy
  for (int rr = blockIdx.y; rr < min(n_root, static_cast<int>(blockIdx.y) + 1); ++rr) {
    const int slot = root_slots[rr];
    const int r = root_index ? root_index[rr] : rr;
    double x[S];
    load_clv<S>(clv + (static_cast<size_t>(slot) * n + site) * (C * S) + c * S, x);
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < S; ++i) t = fma(freqs[part * S + i], x[i], t);
    t = cat_sum<C>(t) * (1.0 / C);
    // the adjoint's seed factor at this root and site, w / (C Ltilde)
    if (seed && c == 0) seed[static_cast<size_t>(r) * n + site] = w != 0.0 ? w * (1.0 / C) / t : 0.0;
    double val = 0.0;
    if (c == 0 && w != 0.0)
      val = w * (log(t) + scl[static_cast<size_t>(slot) * n + site] * LOG_SCALE_THRESHOLD);
    val = wave_sum(val);
    if ((threadIdx.x & 63) == 0) partials[static_cast<size_t>(r) * n_wt + wt] = val;
  }
}

// ----------------------------------------------------------------- mixture
// Per partition: lnL_p = log sum_t P(t) exp(lnL_{p,t}) (AVERAGE) or
// max_t log P(t) + lnL_{p,t} (BEST) over root classes whose log-probability
// passes min_interesting_tree_logprob (ImprovedLoglikelihood.cpp:521-604);
// a max-shifted FP64 log-sum-exp stands in for the reference's MPFR sum.
// Also produces the adjoint seeds beta[p][root CLV] = sum of the class
// weights d lnL_p / d lnL_{p,t}, and d lnL / d p_r.
// The inputs (root sums, class data, d log P / d p) and alpha are staged in
// LDS when they fit in MIX_LDS_DOUBLES (the usual case), so the serial loops
// over classes read LDS instead of making a memory round trip per class;
// otherwise the same code runs on the global arrays.  The arithmetic and
// its order are the same either way.
constexpr int MIX_LDS_DOUBLES = 6144;  // 48 KiB
__host__ __device__ inline int mixture_lds_doubles(int n_root, int nc, int R, int P) {
  return n_root * P + 2 * nc + nc * P + nc * R;
}
__device__ void mixture_body(const DevLayout &L, int n_root, int nc, int R, int variant,
                             const DevClass *__restrict__ classes,
                             const double *__restrict__ class_dlogprob,
                             const double *__restrict__ root_sums, double *__restrict__ beta,
                             double *__restrict__ res, const ResultLayout &rl) {
  __shared__ double best[64], ssum[64];
  __shared__ int arg[64];
  const int P = L.n_parts;  // <= 64 (checked on the host)
  const bool staged = mixture_lds_doubles(n_root, nc, R, P) <= MIX_LDS_DOUBLES;
  // class k: log-probability lp[k], root index ri[k], use flag cu[k]
  const double *rs = root_sums, *cdl = class_dlogprob;
  double *alpha = res + rl.alpha;
  double *lp = nullptr;
  int *ri = nullptr, *cu = nullptr;
  if (staged) {
    double *sm = dyn_lds<double>();
    double *s_rs = sm, *s_lp = s_rs + n_root * P, *s_ri = s_lp + nc, *s_alpha = s_ri + nc, *s_cdl = s_alpha + nc * P;
    for (int i = threadIdx.x; i < n_root * P; i += blockDim.x) s_rs[i] = root_sums[i];
    for (int i = threadIdx.x; i < nc * R; i += blockDim.x) s_cdl[i] = class_dlogprob[i];
    int *s_ri_int = reinterpret_cast<int *>(s_ri), *s_cu = s_ri_int + nc;
    for (int k = threadIdx.x; k < nc; k += blockDim.x) {
      const DevClass cl = classes[k];
      s_lp[k] = cl.logprob;
      s_ri_int[k] = cl.root_index;
      s_cu[k] = cl.use;
    }
    rs = s_rs;
    cdl = s_cdl;
    alpha = s_alpha;
    lp = s_lp;
    ri = s_ri_int;
    cu = s_cu;
    __syncthreads();
  }
  const auto logprob = [&](int k) { return staged ? lp[k] : classes[k].logprob; };
  const auto used = [&](int k) { return (staged ? cu[k] : classes[k].use) != 0; };
  const auto root_of = [&](int k) { return staged ? ri[k] : classes[k].root_index; };
  // per-(root CLV, partition) tree lnL and per-class copies
  for (int idx = threadIdx.x; idx < n_root * P; idx += blockDim.x) {
    const int r = idx / P, p = idx % P;
    res[rl.root_logl + p * n_root + r] = rs[idx];
  }
  for (int idx = threadIdx.x; idx < nc * P; idx += blockDim.x) {
    const int p = idx / nc, k = idx % nc;
    const int r = root_of(k);
    res[rl.tree_logl + p * nc + k] = r >= 0 ? rs[r * P + p] : 0.0;
  }
  // best class per partition (first maximum in class order)
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    double bv = -INFINITY;
    int ba = -1;
    for (int k = 0; k < nc; ++k) {
      if (!used(k)) continue;
      const double v = logprob(k) + rs[root_of(k) * P + p];
      if (v > bv) {
        bv = v;
        ba = k;
      }
    }
    best[p] = bv;
    arg[p] = ba;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < nc * P; idx += blockDim.x) {
    const int p = idx / nc, k = idx % nc;
    double a = 0.0;
    if (variant == 1)
      a = k == arg[p] ? 1.0 : 0.0;
    else if (used(k) && best[p] > -INFINITY)
      a = exp(logprob(k) + rs[root_of(k) * P + p] - best[p]);
    alpha[p * nc + k] = a;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    double sacc = 0.0;
    for (int k = 0; k < nc; ++k) sacc += alpha[p * nc + k];
    ssum[p] = sacc;
    res[rl.per_part + p] = variant == 1 ? best[p] : best[p] + log(sacc);
  }
  __syncthreads();
  if (variant != 1)
    for (int idx = threadIdx.x; idx < nc * P; idx += blockDim.x) {
      const int p = idx / nc;
      alpha[idx] = ssum[p] > 0.0 ? alpha[idx] / ssum[p] : 0.0;
    }
  __syncthreads();
  if (staged)
    for (int idx = threadIdx.x; idx < nc * P; idx += blockDim.x) res[rl.alpha + idx] = alpha[idx];
  for (int idx = threadIdx.x; idx < n_root * P; idx += blockDim.x) {
    const int p = idx / n_root, r = idx % n_root;
    double b = 0.0;
    for (int k = 0; k < nc; ++k)
      if (root_of(k) == r) b += alpha[p * nc + k];
    beta[p * n_root + r] = b;
  }
  for (int r = threadIdx.x; r < R; r += blockDim.x) {
    double g = 0.0;
    for (int p = 0; p < P; ++p)
      for (int k = 0; k < nc; ++k)
        g += alpha[p * nc + k] * cdl[static_cast<size_t>(k) * R + r];
    res[rl.dprob + r] = g;
  }
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int p = 0; p < P; ++p) tot += res[rl.per_part + p];
    res[rl.total] = tot;
  }
}

__global__ void __launch_bounds__(256)
    k_mixture(DevLayout L, int n_root, int nc, int R, int variant,
              const DevClass *__restrict__ classes, const double *__restrict__ class_dlogprob,
              const double *__restrict__ root_sums, double *__restrict__ beta,
              double *__restrict__ res, ResultLayout rl) {
  mixture_body(L, n_root, nc, R, variant, classes, class_dlogprob, root_sums, beta, res, rl);
}

// ------------------------------------------------------- adjoint (FP64)
// Reverse pruning in push form.  For op q with children k (edge e_k, CLV
// x_k, y_k = P_k x_k) and rescale factor F = 2^(256 (k_q - sum k_k)):
//   z_k      = Abar_q * F * prod_{j != k} y_j
//   Abar_k  += P_k^T z_k            (delivered to child k)
//   dJ/dt_k += z_k . (P'_k x_k),    P'_k x_k = r_c Q y_k
// where J = sum_p sum_t beta_{p,t} lnL_{p,t} and Abar is the derivative of
// J w.r.t. the *scaled* CLV values; a root CLV adds the seed
// beta * w_s * (1/C) pi_i / Ltilde_s.  Ops are walked in reverse
// topological order; every consumer of a CLV delivers before the CLV's
// own op runs, the first one writes and later ones accumulate.  The same
// thread owns (site, category) in every op, so accumulation needs neither
// atomics nor barriers and is deterministic.
// One workgroup per tile walks the levels of `walk` (level chunks, top
// level first) and each level's ops in order; since every (site, category)
// has one owning thread in every op, a level's deliveries are read back by
// the thread that stored them, and a single launch can cover any number of
// levels.  The P matrices of each level's first max_staged ops are staged
// in LDS when the walk enters the level.
template <int S>
struct AdjIn {
  double a[S];
  double x[2][S];  // inner children; tips are expanded from tm at use
  double sd;       // seed factor w / (C Ltilde) of a displayed-tree root's site
  int kq, sk[2];   // sk: a child's scaler, or a tip child's code (one register for both)
};

template <int S, int C>
__device__ __forceinline__ void adjoint_inputs(const DevOp &op, int64_t site, int64_t n, int c,
                                               const double *__restrict__ clv,
                                               const int32_t *__restrict__ scl,
                                               const double *__restrict__ adj,
                                               const typename TipType<S>::type *__restrict__ tips,
                                               const double *__restrict__ seed, AdjIn<S> &in) {
  // loads only (defaults for absent inputs are applied at use), so that no
  // load waits for an earlier one
  const size_t row = static_cast<size_t>(C * S);
  in.kq = scl[static_cast<size_t>(op.out) * n + site];
  if (op.flags & OP_HAS_CONSUMERS)
    load_clv<S>(adj + (static_cast<size_t>(op.out) * n + site) * row + c * S, in.a);
  if (op.root >= 0) in.sd = seed[static_cast<size_t>(op.root) * n + site];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int32_t kid = op.kid[k];
    if (k < op.nk && kid >= 0) {
      load_clv<S>(clv + (static_cast<size_t>(kid) * n + site) * row + c * S, in.x[k]);
      in.sk[k] = scl[static_cast<size_t>(kid) * n + site];
    }
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int32_t kid = op.kid[k];
    if (k < op.nk && kid < 0) in.sk[k] = static_cast<int>(tips[static_cast<size_t>(-kid - 1) * n + site]);
  }
}

// At least 3 waves per SIMD (the register count the kernel needs anyway):
// a launch has one wave per 64 (site, category) pairs, about 2.5 per SIMD
// at 41k patterns, so every wave is resident from the start.
template <int S, int C>
__global__ void __launch_bounds__(TILE *C, S == 4 ? 3 : 1)
    k_adjoint_level(const DevOp *__restrict__ ops, const LevelChunk *__restrict__ walk, int n_walk,
                    int max_staged, DevLayout L,
                    const double *__restrict__ clv, const int32_t *__restrict__ scl,
                    double *__restrict__ adj, const typename TipType<S>::type *__restrict__ tips,
                    const double *__restrict__ P, const double *__restrict__ Qd,
                    const double *__restrict__ freqs, const double *__restrict__ seed,
                    const int32_t *__restrict__ tile_part, const double *__restrict__ beta,
                    int n_root, int n_uses, double *__restrict__ grad_partials) {
  // LDS: r_c Q of every category (4 states), then the staged P matrices
  double *sQ = dyn_lds<double>();
  double *sP = sQ + (S == 4 ? C * S * S : 0);
  constexpr int PW = TILE * C / 64;
  const int tile = blockIdx.x;
  if (tile >= L.n_tiles) return;  // grid padded to a multiple of 8 tiles

  const int part = tile_part[tile];
  if constexpr (S == 4) {
    for (int i = threadIdx.x; i < C * S * S; i += blockDim.x)
      sQ[i] = Qd[static_cast<size_t>(part) * C * S * S + i];
    __syncthreads();
  }
  const int c = threadIdx.x % C;
  const int64_t site = static_cast<int64_t>(tile) * TILE + threadIdx.x / C;
  const int64_t n = L.n_sites_pad;
  const size_t row = static_cast<size_t>(C * S);
  const int wave = threadIdx.x / 64;
  const double *Qc = (S == 4 ? sQ : Qd + static_cast<size_t>(part) * C * S * S) + c * (S * S);
  // 4 states: r_c Q of this lane's category in registers (every op reads it;
  // from LDS it would cost as many LDS bytes as a P matrix per child)
  double Qr[S == 4 ? 16 : 1];
  if constexpr (S == 4)
#pragma unroll
    for (int i = 0; i < 16; ++i) Qr[i] = Qc[i];
  for (int w = 0; w < n_walk; ++w) {
    const LevelChunk lv = walk[w];
    const DevOp *__restrict__ lops = ops + lv.op_begin;
    const int n_ops = lv.op_end - lv.op_begin;
    const int n_staged = min(n_ops, max_staged);
    if (S == 4 && w > 0) __syncthreads();  // every lane is done with the previous level's staged P
    stage_ops_p<S, C>(sP, P, lops, n_staged, part, L);  // ends with a barrier (4 states)
    for (int j = 0; j < n_ops; ++j) {
      const DevOp op = lops[j];
      AdjIn<S> cur;
      __builtin_amdgcn_s_setprio(1);  // see the note at wait_loads
      adjoint_inputs<S, C>(op, site, n, c, clv, scl, adj, tips, seed, cur);
      __builtin_amdgcn_s_setprio(0);
      double a[S];
      const bool has_adj = (op.flags & OP_HAS_CONSUMERS) != 0;
#pragma unroll
      for (int i = 0; i < S; ++i) a[i] = has_adj ? cur.a[i] : 0.0;
      int ks = 0;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (k < op.nk) {
          if (op.kid[k] < 0)
            tip_vec<S>(static_cast<uint32_t>(cur.sk[k]), cur.x[k]);
          else
            ks += cur.sk[k];
        }
      }
      if (op.root >= 0) {
        // seed beta * w * (1/C) pi_i / Ltilde at a displayed-tree root
        const double f = beta[part * n_root + op.root] * cur.sd;
#pragma unroll
        for (int i = 0; i < S; ++i) a[i] = fma(f, freqs[part * S + i], a[i]);
      }
      double y[2][S];
      const double F = (cur.kq - ks) > 0 ? SCALE_FACTOR : 1.0;
      // arithmetic of both children first, memory traffic after
      double back[2][S], g[2];
      if constexpr (S == 4) {
        // y_k = P_k x_k, z_0 = a F y_1 (or a F), z_1 = a F y_0, back_k = P_k^T
        // z_k: P_0 is read once (row i gives y_0[i], z_0[i] and row i's share
        // of back_0), P_1 twice -- three matrix reads per op instead of four
        const double *P0 = op_p<S, C>(sP, n_staged, P, part, L, j, 0, op.edge[0], c);
        // (selects, not op.edge[op.nk - 1]: a runtime index would put op on the stack)
        const double *P1 = op.nk == 2 ? op_p<S, C>(sP, n_staged, P, part, L, j, 1, op.edge[1], c) : P0;
        double z[2][S];
        if (op.nk == 2) matvec<S>(P1, cur.x[1], y[1]);
#pragma unroll
        for (int i = 0; i < S; ++i) {
          const double2 r01 = reinterpret_cast<const double2 *>(P0 + i * S)[0];
          const double2 r23 = reinterpret_cast<const double2 *>(P0 + i * S)[1];
          y[0][i] = fma(r23.y, cur.x[0][3], fma(r23.x, cur.x[0][2], fma(r01.y, cur.x[0][1], r01.x * cur.x[0][0])));
          z[0][i] = a[i] * F * (op.nk == 2 ? y[1][i] : 1.0);
          if (i == 0) {
            back[0][0] = r01.x * z[0][0];
            back[0][1] = r01.y * z[0][0];
            back[0][2] = r23.x * z[0][0];
            back[0][3] = r23.y * z[0][0];
          } else {
            back[0][0] = fma(r01.x, z[0][i], back[0][0]);
            back[0][1] = fma(r01.y, z[0][i], back[0][1]);
            back[0][2] = fma(r23.x, z[0][i], back[0][2]);
            back[0][3] = fma(r23.y, z[0][i], back[0][3]);
          }
        }
        if (op.nk == 2) {
#pragma unroll
          for (int i = 0; i < S; ++i) z[1][i] = a[i] * F * y[0][i];
          if (op.kid[1] >= 0) matvec_t<S>(P1, z[1], back[1]);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          if (k < op.nk) {
            double dx[S];
            matvec<S>(Qr, y[k], dx);
            double gk = 0.0;
#pragma unroll
            for (int i = 0; i < S; ++i) gk = fma(z[k][i], dx[i], gk);
            g[k] = wave_sum(gk);
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < 2; ++k)
          if (k < op.nk) matvec<S>(op_p<S, C>(sP, n_staged, P, part, L, j, k, op.edge[k], c), cur.x[k], y[k]);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          if (k < op.nk) {
            double z[S], dx[S];
#pragma unroll
            for (int i = 0; i < S; ++i) z[i] = a[i] * F * (op.nk == 2 ? y[1 - k][i] : 1.0);
            matvec<S>(Qc, y[k], dx);
            double gk = 0.0;
#pragma unroll
            for (int i = 0; i < S; ++i) gk = fma(z[i], dx[i], gk);
            g[k] = wave_sum(gk);
            if (op.kid[k] >= 0)
              matvec_t<S>(op_p<S, C>(sP, n_staged, P, part, L, j, k, op.edge[k], c), z, back[k]);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (k < op.nk) {
          if ((threadIdx.x & 63) == 0) grad_partials[grad_index(tile, n_uses, op.gslot[k], PW, wave)] = g[k];
          const int32_t kid = op.kid[k];
          if (kid >= 0) {
            double *dst = adj + (static_cast<size_t>(kid) * n + site) * row + c * S;
            if (!(op.flags & (k == 0 ? OP_FIRST_KID0 : OP_FIRST_KID1))) {
              double acc[S];
              load_clv<S>(dst, acc);
#pragma unroll
              for (int i = 0; i < S; ++i) back[k][i] += acc[i];
            }
            store_clv<S>(dst, back[k]);
          }
        }
      }
    }
  }
}

// ------------------------------------------- 20 states, FP64 plain form
// One wavefront per rate category: lane l of wave c holds site l of the
// 64-site tile in category c.  Every P (and r_c Q) element a wave reads is
// then the same for all of its lanes, so the matrices come in through the
// scalar data path as SGPR operands of the FMAs -- no vector load and no
// LDS read per matrix element, which is what bounds a (site, category) per
// lane mapping at 20 states (400 elements per matrix).  The per-site
// quantity that spans categories (the scaling maximum) crosses the waves
// through LDS.  The CLV layout is the same as every other FP64 kernel's.
// A matrix the kernel only reads, addressed uniformly across the wave: in
// the constant address space its loads are scalar loads.
using cmat = const __attribute__((address_space(4))) double *;
__device__ __forceinline__ cmat as_cmat(const double *p) { return (cmat)p; }

__device__ __forceinline__ void load_row20(const double *__restrict__ p, double (&x)[20]) {
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const double2 v = reinterpret_cast<const double2 *>(p)[i];
    x[2 * i] = v.x;
    x[2 * i + 1] = v.y;
  }
}

__device__ __forceinline__ void store_row20(double *__restrict__ p, const double (&x)[20]) {
#pragma unroll
  for (int i = 0; i < 10; ++i) reinterpret_cast<double2 *>(p)[i] = make_double2(x[2 * i], x[2 * i + 1]);
}

// Workgroup (tile, op): the op's CLV rows of one 64-site tile, C waves.
template <int C>
__global__ void __launch_bounds__(TILE *C)
    k_forward_level_aa(const DevOp *__restrict__ ops, DevLayout L, double *__restrict__ clv,
                       int32_t *__restrict__ scl, const uint32_t *__restrict__ tips, const double *__restrict__ P,
                       const int32_t *__restrict__ tile_part) {
  constexpr int S = 20;
  __shared__ double smax[C][64];
  const int tile = blockIdx.x;
  if (tile >= L.n_tiles) return;  // grid padded to a multiple of 8 tiles
  const DevOp op = ops[blockIdx.y];
  const int part = tile_part[tile];
  const int c = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int64_t site = static_cast<int64_t>(tile) * TILE + lane;
  const int64_t n = L.n_sites_pad;
  const size_t row = static_cast<size_t>(C * S);
  // loads of both children first; the tip mask and the scaler are loaded
  // for every child (tip 0 and variant 0 stand in for the absent one): a
  // value loaded on one path only would be copied at the join, which waits
  // for every load issued so far
  double x[2][S];
  uint32_t tm[2];
  int sk[2];
  __builtin_amdgcn_s_setprio(1);  // see the note at wait_loads
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int32_t kid = k < op.nk ? op.kid[k] : -1;
    tm[k] = tips[static_cast<size_t>(kid < 0 ? -kid - 1 : 0) * n + site];
    sk[k] = scl[static_cast<size_t>(kid < 0 ? 0 : kid) * n + site];
    if (k < op.nk && kid >= 0) load_row20(clv + (static_cast<size_t>(kid) * n + site) * row + c * S, x[k]);
  }
  __builtin_amdgcn_s_setprio(0);
  int sc = 0;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (k < op.nk) {
      if (op.kid[k] < 0)
        tip_vec<S>(tm[k], x[k]);
      else
        sc += sk[k];
    }
  }
  double v[S];
#pragma unroll
  for (int i = 0; i < S; ++i) v[i] = 1.0;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (k < op.nk) {
      const cmat Pk = as_cmat(P + ((static_cast<size_t>(part) * L.n_edges + op.edge[k]) * C + c) * (S * S));
#pragma unroll
      for (int i = 0; i < S; ++i) {
        double s = 0.0;
#pragma unroll
        for (int jj = 0; jj < S; ++jj) s = fma(Pk[i * S + jj], x[k][jj], s);
        v[i] *= s;
      }
    }
  }
  double m = v[0];
#pragma unroll
  for (int i = 1; i < S; ++i) m = fmax(m, v[i]);
  smax[c][lane] = m;
  __syncthreads();
  double mm = smax[0][lane];
#pragma unroll
  for (int cc = 1; cc < C; ++cc) mm = fmax(mm, smax[cc][lane]);
  if (mm < SCALE_THRESHOLD) {
#pragma unroll
    for (int i = 0; i < S; ++i) v[i] *= SCALE_FACTOR;
    sc += 1;
  }
  store_row20(clv + (static_cast<size_t>(op.out) * n + site) * row + c * S, v);
  if (c == 0) scl[static_cast<size_t>(op.out) * n + site] = sc;
}

// Reverse sweep of one level (push form, as k_adjoint_level): one
// workgroup per tile walks the level's ops in order; per op and child k,
// z_k = Abar F y_other, dJ/dt_k += z_k . (r_c Q y_k), Abar_k += P_k^T z_k,
// formed row by row of P so that z never has to be held whole.
template <int C>
__global__ void __launch_bounds__(TILE *C)
    k_adjoint_level_aa(const DevOp *__restrict__ ops, int n_ops, DevLayout L, const double *__restrict__ clv,
                       const int32_t *__restrict__ scl, double *__restrict__ adj,
                       const uint32_t *__restrict__ tips, const double *__restrict__ P,
                       const double *__restrict__ Qd, const double *__restrict__ freqs,
                       const double *__restrict__ seed, const int32_t *__restrict__ tile_part,
                       const double *__restrict__ beta, int n_root, int n_uses,
                       double *__restrict__ grad_partials) {
  constexpr int S = 20;
  constexpr int PW = TILE * C / 64;
  const int tile = blockIdx.x;
  if (tile >= L.n_tiles) return;  // grid padded to a multiple of 8 tiles
  const int part = tile_part[tile];
  const int c = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int64_t site = static_cast<int64_t>(tile) * TILE + lane;
  const int64_t n = L.n_sites_pad;
  const size_t row = static_cast<size_t>(C * S);
  const double *Qc = Qd + (static_cast<size_t>(part) * C + c) * (S * S);
  // one op of the run (whose ops share no child) per workgroup
  ops += blockIdx.y;
  n_ops = min(n_ops - static_cast<int>(blockIdx.y), 1);
  for (int j = 0; j < n_ops; ++j) {
    const DevOp op = ops[j];
    // load phase: every load issued whatever the op's kind (the adjoint
    // row of an op without consumers, seed row 0, tip 0 and variant 0 stand
    // in for absent inputs and are dropped below), so that no join copies a
    // loaded register and waits for the loads issued before it
    double araw[S], x[2][S];
    __builtin_amdgcn_s_setprio(1);  // see the note at wait_loads
    load_row20(adj + (static_cast<size_t>(op.out) * n + site) * row + c * S, araw);
    const int kq = scl[static_cast<size_t>(op.out) * n + site];
    const double sd = seed[static_cast<size_t>(op.root >= 0 ? op.root : 0) * n + site];
    uint32_t tm[2];
    int sk[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int32_t kid = k < op.nk ? op.kid[k] : -1;
      tm[k] = tips[static_cast<size_t>(kid < 0 ? -kid - 1 : 0) * n + site];
      sk[k] = scl[static_cast<size_t>(kid < 0 ? 0 : kid) * n + site];
      if (k < op.nk && kid >= 0) load_row20(clv + (static_cast<size_t>(kid) * n + site) * row + c * S, x[k]);
    }
    __builtin_amdgcn_s_setprio(0);
    const bool has_adj = (op.flags & OP_HAS_CONSUMERS) != 0;
    double a[S];
#pragma unroll
    for (int i = 0; i < S; ++i) a[i] = has_adj ? araw[i] : 0.0;
    if (op.root >= 0) {
      // seed beta * w * (1/C) pi_i / Ltilde at a displayed-tree root
      const double f = beta[part * n_root + op.root] * sd;
#pragma unroll
      for (int i = 0; i < S; ++i) a[i] = fma(f, freqs[part * S + i], a[i]);
    }
    double y[2][S];
    int ks = 0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (k < op.nk) {
        if (op.kid[k] < 0)
          tip_vec<S>(tm[k], x[k]);
        else
          ks += sk[k];
        const cmat Pk = as_cmat(P + ((static_cast<size_t>(part) * L.n_edges + op.edge[k]) * C + c) * (S * S));
#pragma unroll
        for (int i = 0; i < S; ++i) {
          double s = 0.0;
#pragma unroll
          for (int jj = 0; jj < S; ++jj) s = fma(Pk[i * S + jj], x[k][jj], s);
          y[k][i] = s;
        }
      }
    }
    const double F = (kq - ks) > 0 ? SCALE_FACTOR : 1.0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (k >= op.nk) continue;
      const int32_t kid = op.kid[k];
      cmat Pk = as_cmat(P + ((static_cast<size_t>(part) * L.n_edges + op.edge[k]) * C + c) * (S * S));
      cmat Qk = as_cmat(Qc);
      // opaque to the optimiser: the matrices are read again here, through
      // the scalar cache, instead of being kept in (or hoisted out of the op
      // loop into) 800 SGPRs each, which would spill
      asm volatile("" : "+s"(Pk));
      asm volatile("" : "+s"(Qk));
      double back[S], gk = 0.0;
      // one straight-line body per case (internal child or tip): a branch
      // inside the row loop splits it into blocks across which the scalar
      // registers holding matrix rows stay live and spill
      auto rows = [&](auto with_back) {
#pragma unroll
        for (int i = 0; i < S; ++i) {
          const double zi = a[i] * F * (op.nk == 2 ? y[1 - k][i] : 1.0);
          double d = 0.0;
#pragma unroll
          for (int jj = 0; jj < S; ++jj) d = fma(Qk[i * S + jj], y[k][jj], d);
          gk = fma(zi, d, gk);
          if constexpr (decltype(with_back)::value) {
#pragma unroll
            for (int jj = 0; jj < S; ++jj) back[jj] = i == 0 ? Pk[jj] * zi : fma(Pk[i * S + jj], zi, back[jj]);
          }
        }
      };
      if (kid >= 0)
        rows(std::true_type{});
      else
        rows(std::false_type{});
      const double g = wave_sum(gk);
      if (lane == 0) grad_partials[grad_index(tile, n_uses, op.gslot[k], PW, c)] = g;
      if (kid >= 0) {
        double *dst = adj + (static_cast<size_t>(kid) * n + site) * row + c * S;
        if (!(op.flags & (k == 0 ? OP_FIRST_KID0 : OP_FIRST_KID1))) {
          double acc[S];
          load_row20(dst, acc);
#pragma unroll
          for (int i = 0; i < S; ++i) back[i] += acc[i];
        }
        store_row20(dst, back);
      }
    }
  }
}

// ---------------------------------------------- model gradient (FP64)
// dlnL / dP_{e,c} for every (op, child) use: G = sum over sites of z (x)^T
// with z the adjoint of y = P x (recomputed from the op's stored adjoint,
// its children and the displayed-tree seed, exactly as k_adjoint_level
// forms it) and x the child's CLV; and, at displayed-tree roots, the direct
// dependence of lnL on the frequencies, sum_s beta w (1/C) x_i / Ltilde.
// Grid (op, chunk): a chunk is a 1/n_split slice of one partition's tiles,
// so every partial belongs to one partition; a workgroup walks its tiles
// and reduces its threads' sums once, in a fixed order.
template <int S, int C>
__global__ void __launch_bounds__(TILE *C)
    k_model_grad(const DevOp *__restrict__ ops, DevLayout L, const double *__restrict__ clv,
                 const int32_t *__restrict__ scl, const double *__restrict__ adj,
                 const typename TipType<S>::type *__restrict__ tips, const double *__restrict__ P,
                 const double *__restrict__ freqs, const double *__restrict__ weights,
                 const double *__restrict__ beta, int n_root, const int32_t *__restrict__ part_tile_begin,
                 int n_split, int n_uses, double *__restrict__ g_out, double *__restrict__ r_out) {
  __shared__ double red[TILE * C / 64][2 * S * S + S];
  const DevOp op = ops[blockIdx.x];
  const int chunk = blockIdx.y;
  const int part = chunk / n_split, slice = chunk % n_split;
  const int tb = part_tile_begin[part], te = part_tile_begin[part + 1];
  const int t0 = tb + (te - tb) * slice / n_split, t1 = tb + (te - tb) * (slice + 1) / n_split;
  const int c = threadIdx.x % C;
  const int64_t n = L.n_sites_pad;
  const size_t row = static_cast<size_t>(C * S);
  const double *Pk[2];
#pragma unroll
  for (int k = 0; k < 2; ++k)
    Pk[k] = P + ((static_cast<size_t>(part) * L.n_edges + (k < op.nk ? op.edge[k] : 0)) * C + c) * (S * S);
  double g[2][S * S], rr[S];
#pragma unroll
  for (int q = 0; q < S * S; ++q) g[0][q] = g[1][q] = 0.0;
#pragma unroll
  for (int i = 0; i < S; ++i) rr[i] = 0.0;
  const double bt = op.root >= 0 ? beta[part * n_root + op.root] : 0.0;
  for (int tile = t0; tile < t1; ++tile) {
    const int64_t site = static_cast<int64_t>(tile) * TILE + threadIdx.x / C;
    const double w = weights[site];
    double a[S], x[2][S], y[2][S];
#pragma unroll
    for (int i = 0; i < S; ++i) a[i] = 0.0;
    if (op.flags & OP_HAS_CONSUMERS) load_clv<S>(adj + (static_cast<size_t>(op.out) * n + site) * row + c * S, a);
    int ks = 0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (k >= op.nk) continue;
      const int32_t kid = op.kid[k];
      if (kid >= 0) {
        load_clv<S>(clv + (static_cast<size_t>(kid) * n + site) * row + c * S, x[k]);
        ks += scl[static_cast<size_t>(kid) * n + site];
      } else {
        tip_vec<S>(tips[static_cast<size_t>(-kid - 1) * n + site], x[k]);
      }
      matvec<S>(Pk[k], x[k], y[k]);
    }
    if (op.root >= 0) {
      // seed beta w (1/C) pi_i / Ltilde of a displayed-tree root, and the
      // frequencies' direct share beta w (1/C) x_i / Ltilde
      double xr[S];
      load_clv<S>(clv + (static_cast<size_t>(op.out) * n + site) * row + c * S, xr);
      double t = 0.0;
#pragma unroll
      for (int i = 0; i < S; ++i) t = fma(freqs[part * S + i], xr[i], t);
      t = cat_sum<C>(t) * (1.0 / C);
      const double f = w != 0.0 ? bt * w * (1.0 / C) / t : 0.0;
#pragma unroll
      for (int i = 0; i < S; ++i) {
        a[i] = fma(f, freqs[part * S + i], a[i]);
        rr[i] = fma(f, xr[i], rr[i]);
      }
    }
    const double F = (scl[static_cast<size_t>(op.out) * n + site] - ks) > 0 ? SCALE_FACTOR : 1.0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (k >= op.nk) continue;
#pragma unroll
      for (int i = 0; i < S; ++i) {
        const double z = a[i] * F * (op.nk == 2 ? y[1 - k][i] : 1.0);
#pragma unroll
        for (int j = 0; j < S; ++j) g[k][i * S + j] = fma(z, x[k][j], g[k][i * S + j]);
      }
    }
  }
  // sums over the sites of each category: lanes of one category are C apart
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  auto site_sum = [&](double v) {
#pragma unroll
    for (int off = C; off < 64; off <<= 1) v += __shfl_xor(v, off);
    return v;
  };
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int q = 0; q < S * S; ++q) g[k][q] = site_sum(g[k][q]);
  // lane c (< C) of every wave now holds its category's wave sums; combine
  // the waves through LDS in wave order, one category at a time
  for (int cc = 0; cc < C; ++cc) {
    if (lane == cc) {
#pragma unroll
      for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int q = 0; q < S * S; ++q) red[wave][k * S * S + q] = g[k][q];
    }
    __syncthreads();
    if (threadIdx.x < 2 * S * S) {
      const int k = threadIdx.x / (S * S), q = threadIdx.x % (S * S);
      double sum = 0.0;
      for (int wv = 0; wv < TILE * C / 64; ++wv) sum += red[wv][threadIdx.x];
      if (k < op.nk)
        g_out[((static_cast<size_t>(op.gslot[k]) * gridDim.y + chunk) * C + cc) * (S * S) + q] = sum;
    }
    __syncthreads();
  }
  if (op.root >= 0) {
#pragma unroll
    for (int i = 0; i < S; ++i) {
      double v = rr[i];
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off);
      rr[i] = v;
    }
    if (lane == 0)
#pragma unroll
      for (int i = 0; i < S; ++i) red[wave][2 * S * S + i] = rr[i];
    __syncthreads();
    if (threadIdx.x < S) {
      double sum = 0.0;
      for (int wv = 0; wv < TILE * C / 64; ++wv) sum += red[wv][2 * S * S + threadIdx.x];
      r_out[(static_cast<size_t>(op.root) * gridDim.y + chunk) * S + threadIdx.x] = sum;
    }
  }
}

// ------------------------------------------------------- reductions
// out[row][p] = sum of in[row][t] over the partials t of partition p
// (partials_per_tile per tile, tiles of a partition are contiguous).  One
// workgroup per (row, partition): coalesced strided loads, fixed tree
// order, so the sums are deterministic.
__global__ void __launch_bounds__(256)
    k_reduce_segments(const double *__restrict__ in, size_t row_stride, int pw,
                      const int32_t *__restrict__ part_tile_begin, int P,
                      double *__restrict__ out) {
  __shared__ double sh[4];
  const int row = blockIdx.x / P, p = blockIdx.x % P;
  const double *r = in + static_cast<size_t>(row) * row_stride;
  const int t0 = part_tile_begin[p] * pw, t1 = part_tile_begin[p + 1] * pw;
  double s = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += 256) s += r[t];
  s = block_sum<256>(s, sh);
  if (threadIdx.x == 0) out[static_cast<size_t>(row) * P + p] = s;
}

// part_dbrlen[p][e] = sum over the uses of edge e of use_sums[u][p];
// dbrlen[e] = sum over partitions.
__global__ void k_grad_total(DevLayout L, const int32_t *__restrict__ edge_use_begin,
                             const int32_t *__restrict__ edge_use_list,
                             const double *__restrict__ use_sums, double *__restrict__ res,
                             ResultLayout rl) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L.n_edges) return;
  const int P = L.n_parts;
  double tot = 0.0;
  for (int p = 0; p < P; ++p) {
    double s = 0.0;
    for (int i = edge_use_begin[e]; i < edge_use_begin[e + 1]; ++i)
      s += use_sums[static_cast<size_t>(edge_use_list[i]) * P + p];
    res[rl.part_dbrlen + p * L.n_edges + e] = s;
    tot += s;
  }
  res[rl.dbrlen + e] = tot;
}

// ==================================================== FP32 batched form
// One workgroup per 64-site tile walks the whole op list (children before
// parents, depth-first post-order so a child is consumed soon after it is
// written).  Thread (site, category) owns its CLV row in every op, so every
// value a thread reads was written by the same thread: no barrier is needed
// after the P tables are staged in LDS.  CLVs are FP32, renormalised at
// every op by an exact power of two so the site maximum lies in [0.5, 1);
// the per-site exponent k (true = stored * 2^-k) is carried as an int32.
// Site sums, the mixture and the gradient reduction across waves are FP64;
// inside a wave the gradient's per-(site, category) shares (each of order
// weight / branch length) are summed in FP32 per row of 16 lanes.
constexpr double LN2 = 0.69314718055994530941723212145818;
constexpr int EXP_EMPTY = -(1 << 28);  // exponent of an all-zero adjoint

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

// Maximum over the C category lanes of a site, comparing the float bit
// patterns as signed integers: the float order for values >= 0, and every
// positive value beats every negative one, which is all a site maximum
// needs.  (fmaxf would add a NaN canonicalisation around every cross-lane
// move; the integer max takes the DPP move as an operand modifier.)
template <int C>
__device__ __forceinline__ float cat_maxf(float m) {
  int u = __float_as_int(m);
  if constexpr (C > 1) u = max(u, dpp_i<DPP_XOR1>(u));
  if constexpr (C > 2) u = max(u, dpp_i<DPP_XOR2>(u));
  if constexpr (C > 4) u = max(u, __shfl_xor(u, 4));
  return __int_as_float(u);
}

// Sum over the 64 lanes of a wave of FP32 values, in FP32 and a fixed order,
// returned in FP64: four row_shr steps inside each row of 16 lanes (lane 15
// of a row ends with the row's sum S_r), then row_bcast:15 (0x142: each row
// adds lane 15 of the row before it, so lane 31 = S0 + S1 and lane 63 =
// S2 + S3) and row_bcast:31 (0x143: rows 2 and 3 add lane 31), after which
// lane 63 holds the total; one lane read instead of four.
__device__ __forceinline__ double wave_sum_f32(float v) {
  v += dpp<0x111>(v);
  v += dpp<0x112>(v);
  v += dpp<0x114>(v);
  v += dpp<0x118>(v);
  v += dpp<0x142>(v);
  v += dpp<0x143>(v);
  return static_cast<double>(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)));
}

__device__ __forceinline__ void tip_vec4(uint32_t mask, float (&x)[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) x[i] = ((mask >> i) & 1u) ? 1.0f : 0.0f;
}

__device__ __forceinline__ void load4(const float *__restrict__ p, float (&x)[4]) {
  const float4 v = *reinterpret_cast<const float4 *>(p);
  x[0] = v.x;
  x[1] = v.y;
  x[2] = v.z;
  x[3] = v.w;
}

__device__ __forceinline__ void store4(float *__restrict__ p, const float (&x)[4]) {
  *reinterpret_cast<float4 *>(p) = make_float4(x[0], x[1], x[2], x[3]);
}

#ifdef __HIP_DEVICE_COMPILE__
__device__ __forceinline__ void load4(const GAS float *__restrict__ p, float (&x)[4]) {
  const float4 v = *reinterpret_cast<const GAS float4 *>(p);
  x[0] = v.x;
  x[1] = v.y;
  x[2] = v.z;
  x[3] = v.w;
}

__device__ __forceinline__ void store4(GAS float *__restrict__ p, const float (&x)[4]) {
  *reinterpret_cast<GAS float4 *>(p) = make_float4(x[0], x[1], x[2], x[3]);
}
#endif

__device__ __forceinline__ void mv4(const float *__restrict__ M, const float (&x)[4], float (&y)[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
    y[i] = fmaf(M[4 * i + 3], x[3], fmaf(M[4 * i + 2], x[2], fmaf(M[4 * i + 1], x[1], M[4 * i] * x[0])));
}

__device__ __forceinline__ void mv4t(const float *__restrict__ M, const float (&x)[4], float (&y)[4]) {
#pragma unroll
  for (int j = 0; j < 4; ++j)
    y[j] = fmaf(M[12 + j], x[3], fmaf(M[8 + j], x[2], fmaf(M[4 + j], x[1], M[j] * x[0])));
}

// Raw inputs of one group of the 4-state FP32 forward: child rows, child
// scaler words and tip masks (expanded when used).
template <int G>
struct FwdIn {
  float x[G][2][4];
  int sk[G][2];
  uint32_t tm[G][2];
};

// Issues every load of group [g0, g0 + nb): inner children in one pass, tip
// masks in another (no join of a load path with a non-load path between two
// loads), nothing waits.
template <int C, int G>
__device__ __forceinline__ void fwd_load(const CAS DevOp *__restrict__ ops, int g0, int nb,
                                         const GAS float *__restrict__ clv,
                                         const GAS int32_t *__restrict__ scl,
                                         const GAS uint8_t *__restrict__ tips, int64_t n,
                                         int64_t site, int c, FwdIn<G> &in) {
#pragma unroll
  for (int j = 0; j < G; ++j) {
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      if (j < nb) {
        const DevOp op = ops[g0 + j];
        const int32_t kid = op.kid[kk];
        if (kk < op.nk && kid >= 0) {
          load4(clv + (static_cast<size_t>(kid) * n + site) * (C * 4) + c * 4, in.x[j][kk]);
          in.sk[j][kk] = scl[static_cast<size_t>(kid) * n + site];
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < G; ++j) {
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      if (j < nb) {
        const DevOp op = ops[g0 + j];
        const int32_t kid = op.kid[kk];
        if (kk < op.nk && kid < 0) in.tm[j][kk] = tips[static_cast<size_t>(-kid - 1) * n + site];
      }
    }
  }
}

// TS sites per workgroup (TS C threads, a multiple of 64, TS <= 64 so a
// workgroup never straddles partitions) and G ops per group are the
// autotuned shape parameters.
// The bodies of the persistent kernels take every buffer as a __restrict__
// parameter: restrict only reaches the optimiser through parameters, and it
// lets the compiler keep op-table loads scalar and move CLV loads past
// stores to other buffers.
template <int C, int G, int TS>
__device__ __forceinline__ void forward_persistent_body(
    const CAS DevOp *__restrict__ ops, const CAS int32_t *__restrict__ group_begin, int n_groups,
    const DevLayout L, GAS float *__restrict__ clv, GAS int32_t *__restrict__ scl,
    const GAS uint8_t *__restrict__ tips, const GAS float *__restrict__ Pf,
    const GAS int32_t *__restrict__ tile_part, const GAS double *__restrict__ freqs,
    const GAS double *__restrict__ weights, GAS double *__restrict__ root_partials, int tile) {
  float *sP = dyn_lds<float>();  // P of this tile's partition: [edge][cat][4][4]
  constexpr int PW = TILE * C / 64;  // site-sum partials (waves) per 64-site tile
  const int64_t site0 = static_cast<int64_t>(tile) * TS;
  const int part = tile_part[site0 / TILE];
  const int np = L.n_edges * C * 16;
  const GAS float *gP = Pf + static_cast<size_t>(part) * np;
  for (int i = threadIdx.x; i < np; i += blockDim.x) sP[i] = gP[i];
  __syncthreads();
  const int c = threadIdx.x % C;
  const int64_t site = site0 + threadIdx.x / C;
  const int64_t n = L.n_sites_pad;
  const size_t n_wt = static_cast<size_t>(L.n_tiles) * PW;
  const size_t wt = static_cast<size_t>(site0 * C / 64) + threadIdx.x / 64;  // global wave
  const double w = weights[site];
  double pi[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) pi[i] = freqs[part * 4 + i];
  // The inputs of a group are loaded as raw values (rows, scaler words, tip
  // masks) in one phase with no wait inside it.  When the next group does
  // not consume this one (OP_GROUP_INDEP on its first op), its loads are
  // issued before this group's arithmetic and stores, so the two groups'
  // memory round trips overlap.
  FwdIn<G> cur, nxt;
  fwd_load<C, G>(ops, group_begin[0], group_begin[1] - group_begin[0], clv, scl, tips, n, site, c, cur);
  wait_loads();
  for (int gi = 0; gi < n_groups; ++gi) {
    const int g0 = group_begin[gi];
    const int nb = group_begin[gi + 1] - g0;
    const bool has_next = gi + 1 < n_groups;
    int n0 = 0, nn = 0;
    bool early = false;
    if (has_next) {
      n0 = group_begin[gi + 1];
      nn = group_begin[gi + 2] - n0;
      early = (ops[n0].flags & OP_GROUP_INDEP) != 0;
    }
    if (early) {
      __builtin_amdgcn_s_setprio(1);  // see the note at wait_loads
      fwd_load<C, G>(ops, n0, nn, clv, scl, tips, n, site, c, nxt);
      __builtin_amdgcn_s_setprio(0);
    }
    float v[G][4];
    int kk_[G];
    double rv[G];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      if (j < nb) {
        const DevOp op = ops[g0 + j];
        int ks = 0;
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
          if (kk < op.nk) {
            if (op.kid[kk] < 0)
              tip_vec4(cur.tm[j][kk], cur.x[j][kk]);
            else
              ks += cur.sk[j][kk];
          }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) v[j][i] = 1.f;
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
          if (kk < op.nk) {
            float y[4];
            mv4(sP + (op.edge[kk] * C + c) * 16, cur.x[j][kk], y);
#pragma unroll
            for (int i = 0; i < 4; ++i) v[j][i] *= y[i];
          }
        }
        const float m = cat_maxf<C>(fmaxf(fmaxf(v[j][0], v[j][1]), fmaxf(v[j][2], v[j][3])));
        int e = 0;
        if (m > 0.f) (void)frexpf(m, &e);
#pragma unroll
        for (int i = 0; i < 4; ++i) v[j][i] = ldexpf(v[j][i], -e);
        kk_[j] = ks - e;
        rv[j] = 0.0;
        if (op.root >= 0) {
          double t = 0.0;
#pragma unroll
          for (int i = 0; i < 4; ++i) t = fma(pi[i], static_cast<double>(v[j][i]), t);
          t = cat_sum<C>(t) * (1.0 / C);
          const double val = (c == 0 && w != 0.0) ? w * (log(t) - kk_[j] * LN2) : 0.0;
          rv[j] = wave_sum(val);
        }
      }
    }
    if (early) wait_loads();  // the next group's inputs, before this group's stores
#pragma unroll
    for (int j = 0; j < G; ++j) {
      if (j < nb) {
        const DevOp op = ops[g0 + j];
        store4(clv + (static_cast<size_t>(op.out) * n + site) * (C * 4) + c * 4, v[j]);
        if (c == 0) scl[static_cast<size_t>(op.out) * n + site] = kk_[j];
        if (op.root >= 0 && (threadIdx.x & 63) == 0)
          root_partials[static_cast<size_t>(op.root) * n_wt + wt] = rv[j];
      }
    }
    if (has_next) {
      if (!early) {
        __builtin_amdgcn_s_setprio(1);
        fwd_load<C, G>(ops, n0, nn, clv, scl, tips, n, site, c, nxt);
        __builtin_amdgcn_s_setprio(0);
        wait_loads();
      }
      cur = nxt;
    }
  }
}

template <int C, int G, int TS>
__global__ void __launch_bounds__(TS *C)
    k_forward_persistent(const NetDesc *__restrict__ nets, const WorkItem *__restrict__ work) {
  const WorkItem wi = work[blockIdx.x];
  const NetDesc &nd = nets[wi.net];
  forward_persistent_body<C, G, TS>(to_constant(nd.ops), to_constant(nd.group_begin), nd.n_groups, nd.L,
                                    to_global(nd.clv), to_global(nd.scl), to_global(nd.tips),
                                    to_global(nd.Pf), to_global(nd.tile_part), to_global(nd.freqs),
                                    to_global(nd.weights), to_global(nd.root_partials), wi.tile);
}

// Push-form reverse sweep (as k_adjoint_level) over the groups of the
// forward order, backwards.  The adjoint of every CLV is stored as FP32
// mantissas with a per-site exponent g: dJ/d(stored CLV) = stored * 2^g,
// renormalised at every write; contributions are aligned to the larger
// exponent before they are added, so no intermediate can overflow.
// Site maximum of |x| over the C category lanes of a site.
template <int C>
__device__ __forceinline__ float site_absmax(const float (&x)[4]) {
  return cat_maxf<C>(fmaxf(fmaxf(fabsf(x[0]), fabsf(x[1])), fmaxf(fabsf(x[2]), fabsf(x[3]))));
}

// x 2^ex + y 2^ey with both aligned to the larger exponent, which is
// returned.  Branch-free (one side shifts by zero): the comparison differs
// between the sites of a wave.
__device__ __forceinline__ int align_add(float (&x)[4], int ex, const float (&y)[4], int ey) {
  const int e = max(ex, ey);
  // power-of-two scales: the products are exact, as ldexp would be
  const float sx = ldexpf(1.0f, ex - e), sy = ldexpf(1.0f, ey - e);
#pragma unroll
  for (int i = 0; i < 4; ++i) x[i] = fmaf(x[i], sx, y[i] * sy);
  return e;
}

// Scales x so that its site maximum m of |x| lands in [0.5, 1) and returns
// the new exponent; EXP_EMPTY for an all-zero site (frexp of 0 gives 0, so
// the scaling is then the identity).
__device__ __forceinline__ int renorm(float (&x)[4], int ex, float m) {
  int e = 0;
  (void)frexpf(m, &e);
#pragma unroll
  for (int i = 0; i < 4; ++i) x[i] = ldexpf(x[i], -e);
  return m > 0.f ? ex + e : EXP_EMPTY;
}

// Adds delivered vector b (adjoint exponent eb) to the accumulated value
// (acc, ea) unless this is the child's first delivery, renormalises and
// stores.  An all-zero delivery takes exponent EXP_EMPTY so that it cannot
// scale the accumulated value away.
template <int C>
__device__ __forceinline__ void finish_delivery(GAS float *__restrict__ dst_m, GAS int32_t *__restrict__ dst_e,
                                                bool first, bool lane0, float (&b)[4], int eb,
                                                const float (&acc)[4], int ea) {
  float mb = site_absmax<C>(b);
  if (!first) {  // uniform: a flag of the op
    eb = align_add(b, mb > 0.f ? eb : EXP_EMPTY, acc, ea);
    mb = site_absmax<C>(b);
  }
  eb = renorm(b, eb, mb);
  store4(dst_m, b);
  if (lane0) *dst_e = eb;
}

// A delivery to a child whose adjoint lives in an LDS pass slot (see
// OP_PASS_KID0): the same arithmetic as finish_delivery, on the thread's
// slot instead of HBM.  slot_e is the site's exponent, shared by the C
// lanes of the site (each writes the same value).
template <int C>
__device__ __forceinline__ void pass_delivery(float *__restrict__ slot_m, int *__restrict__ slot_e, bool first,
                                              float (&b)[4], int eb) {
  float mb = site_absmax<C>(b);
  if (!first) {  // uniform: a flag of the op
    const float4 v = *reinterpret_cast<const float4 *>(slot_m);
    const float acc[4] = {v.x, v.y, v.z, v.w};
    eb = align_add(b, mb > 0.f ? eb : EXP_EMPTY, acc, *slot_e);
    mb = site_absmax<C>(b);
  }
  eb = renorm(b, eb, mb);
  *reinterpret_cast<float4 *>(slot_m) = make_float4(b[0], b[1], b[2], b[3]);
  *slot_e = eb;
}

template <int C, int G, int TS>
__device__ __forceinline__ void adjoint_persistent_body(
    const CAS DevOp *__restrict__ ops, const CAS int32_t *__restrict__ group_begin, int n_groups,
    const DevLayout L, const GAS float *__restrict__ clv, const GAS int32_t *__restrict__ scl,
    GAS float *__restrict__ adj, GAS int32_t *__restrict__ adj_exp,
    const GAS uint8_t *__restrict__ tips, const GAS float *__restrict__ Pf,
    const GAS float *__restrict__ Qf, const GAS int32_t *__restrict__ tile_part,
    const GAS double *__restrict__ freqs, const GAS double *__restrict__ weights,
    const GAS double *__restrict__ beta, int n_root, int n_uses, int n_pass,
    GAS double *__restrict__ grad_partials, int tile) {
  float *sP = dyn_lds<float>();  // [edge][cat][16] then Q: [cat][16], then the pass slots
  constexpr int PWB = TS * C / 64;  // waves of this site block
  constexpr int NT = TS * C;
  const int64_t site0 = static_cast<int64_t>(tile) * TS;
  const int part = tile_part[site0 / TILE];
  const int np = L.n_edges * C * 16;
  const GAS float *gP = Pf + static_cast<size_t>(part) * np;
  for (int i = threadIdx.x; i < np; i += blockDim.x) sP[i] = gP[i];
  for (int i = threadIdx.x; i < C * 16; i += blockDim.x) sP[np + i] = Qf[part * C * 16 + i];
  __syncthreads();
  const int c = threadIdx.x % C;
  // r_c Q of this lane's category, kept in registers (every op reads it)
  float Qr[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) Qr[i] = sP[np + c * 16 + i];
  // pass slots: mantissas [slot][thread][4], site exponents [slot][site]
  float *pass_m = sP + ((np + C * 16 + 3) & ~3);
  int *pass_e = reinterpret_cast<int *>(pass_m + n_pass * NT * 4);
  const int site_in_tile = threadIdx.x / C;
  const int64_t site = site0 + threadIdx.x / C;
  const int64_t n = L.n_sites_pad;
  const int wave = threadIdx.x / 64;
  const bool lane0 = c == 0;
  const double w = weights[site];
  double pi[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) pi[i] = freqs[part * 4 + i];
  const int64_t row_elems = n * C * 4;  // floats of one CLV / adjoint slot
  const uint32_t off_row = static_cast<uint32_t>((site * C + c) * 4) * 4u;  // bytes into such a slot
  const uint32_t off_site = static_cast<uint32_t>(site) * 4u;  // bytes into a scaler / exponent slot
  for (int gi = n_groups - 1; gi >= 0; --gi) {
    const int g0 = group_begin[gi];
    const int nb = group_begin[gi + 1] - g0;
    // load the adjoints and children of the whole group first, as raw
    // values: nothing in this phase waits for an earlier load
    __builtin_amdgcn_s_setprio(1);
    float aq[G][4], x[G][2][4], xq[G][4], acc[G][2][4];
    // sk: a child's scaler, or a tip child's code (one register for both)
    int gq[G], kq[G], sk[G][2], eacc[G][2];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      if (j < nb) {
        const DevOp op = ops[g0 + j];
        kq[j] = *at(scl + op.out * n, off_site);
        if (op.flags & OP_HAS_CONSUMERS) {
          if (op.flags & OP_A_PASS) {
            const int slot = pass_slot(op.flags, 2);
            const float4 v = reinterpret_cast<const float4 *>(pass_m)[slot * NT + threadIdx.x];
            aq[j][0] = v.x;
            aq[j][1] = v.y;
            aq[j][2] = v.z;
            aq[j][3] = v.w;
            gq[j] = pass_e[slot * TS + site_in_tile];
          } else {
            load4(at(adj + op.out * row_elems, off_row), aq[j]);
            gq[j] = *at(adj_exp + op.out * n, off_site);
          }
        }
        if (op.root >= 0) load4(at(clv + op.out * row_elems, off_row), xq[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        if (j < nb) {
          const DevOp op = ops[g0 + j];
          const int32_t kid = op.kid[kk];
          if (kk < op.nk && kid >= 0) {
            load4(at(clv + kid * row_elems, off_row), x[j][kk]);
            sk[j][kk] = *at(scl + kid * n, off_site);
            if (op.flags & (kk == 0 ? OP_ACC_EARLY0 : OP_ACC_EARLY1)) {
              // the child's accumulated adjoint, last written two or more groups ago
              load4(at(adj + kid * row_elems, off_row), acc[j][kk]);
              eacc[j][kk] = *at(adj_exp + kid * n, off_site);
            }
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        if (j < nb) {
          const DevOp op = ops[g0 + j];
          const int32_t kid = op.kid[kk];
          if (kk < op.nk && kid < 0) sk[j][kk] = *at(tips + (-kid - 1) * n, static_cast<uint32_t>(site));
        }
      }
    }
    __builtin_amdgcn_s_setprio(0);
    wait_loads();
    int ks[G];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      if (j < nb) {
        const DevOp op = ops[g0 + j];
        ks[j] = 0;
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
          if (kk < op.nk) {
            if (op.kid[kk] < 0)
              tip_vec4(static_cast<uint32_t>(sk[j][kk]), x[j][kk]);
            else
              ks[j] += sk[j][kk];
          }
        }
        const bool has_adj = (op.flags & OP_HAS_CONSUMERS) != 0;
        float a[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = has_adj ? aq[j][i] : 0.f;
        int ea = has_adj ? gq[j] : EXP_EMPTY;
        if (op.root >= 0) {
          double t = 0.0;
#pragma unroll
          for (int i = 0; i < 4; ++i) t = fma(pi[i], static_cast<double>(xq[j][i]), t);
          t = cat_sum<C>(t) * (1.0 / C);
          const double b = beta[part * n_root + op.root] * w;
          int ef = EXP_EMPTY;
          float sd[4] = {0.f, 0.f, 0.f, 0.f};
          if (b != 0.0) {
            const double mf = frexp(b / (C * t), &ef);
#pragma unroll
            for (int i = 0; i < 4; ++i) sd[i] = static_cast<float>(mf * pi[i]);
          }
          ea = align_add(a, ea, sd, ef);
        }
        // y_k = P_k x_k, z_0 = a y_1 (or a), z_1 = a y_0 and back_k = P_k^T z_k
        // read P_0 once (row i gives y_0[i], then z_0[i] and row i's share of
        // back_0) and P_1 twice: three matrix reads from LDS per op, not four
        float y[2][4], z[2][4], back[2][4];
        const float *P0 = sP + (op.edge[0] * C + c) * 16;
        // (a select, not op.edge[op.nk - 1]: a runtime index would put op on the stack)
        const float *P1 = sP + ((op.nk == 2 ? op.edge[1] : op.edge[0]) * C + c) * 16;
        if (op.nk == 2) mv4(P1, x[j][1], y[1]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float4 r = reinterpret_cast<const float4 *>(P0)[i];
          y[0][i] = fmaf(r.w, x[j][0][3], fmaf(r.z, x[j][0][2], fmaf(r.y, x[j][0][1], r.x * x[j][0][0])));
          z[0][i] = op.nk == 2 ? a[i] * y[1][i] : a[i];
          if (i == 0) {
            back[0][0] = r.x * z[0][0];
            back[0][1] = r.y * z[0][0];
            back[0][2] = r.z * z[0][0];
            back[0][3] = r.w * z[0][0];
          } else {
            back[0][0] = fmaf(r.x, z[0][i], back[0][0]);
            back[0][1] = fmaf(r.y, z[0][i], back[0][1]);
            back[0][2] = fmaf(r.z, z[0][i], back[0][2]);
            back[0][3] = fmaf(r.w, z[0][i], back[0][3]);
          }
        }
        if (op.nk == 2) {
#pragma unroll
          for (int i = 0; i < 4; ++i) z[1][i] = a[i] * y[0][i];
          mv4t(P1, z[1], back[1]);
        }
        const int ez = ea + kq[j] - ks[j];
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
          if (kk < op.nk) {
            float qx[4];
            mv4(Qr, y[kk], qx);
            const float gd =
                fmaf(z[kk][3], qx[3], fmaf(z[kk][2], qx[2], fmaf(z[kk][1], qx[1], z[kk][0] * qx[0])));
            // the site's share of dlnL/dt is O(weight / t): FP32 holds it
            const double gv = wave_sum_f32(ldexpf(gd, ez));
            if ((threadIdx.x & 63) == 0) grad_partials[grad_index(tile, n_uses, op.gslot[kk], PWB, wave)] = gv;
            const int32_t kid = op.kid[kk];
            if (kid >= 0 && (op.flags & (kk == 0 ? OP_PASS_KID0 : OP_PASS_KID1))) {
              const int slot = pass_slot(op.flags, kk);
              pass_delivery<C>(pass_m + 4 * (slot * NT + threadIdx.x), pass_e + slot * TS + site_in_tile,
                               (op.flags & (kk == 0 ? OP_FIRST_KID0 : OP_FIRST_KID1)) != 0, back[kk], ez);
            } else if (kid >= 0) {
              GAS float *dst_m = at(adj + kid * row_elems, off_row);
              GAS int32_t *dst_e = at(adj_exp + kid * n, off_site);
              const bool first = (op.flags & (kk == 0 ? OP_FIRST_KID0 : OP_FIRST_KID1)) != 0;
              if (first || (op.flags & (kk == 0 ? OP_ACC_EARLY0 : OP_ACC_EARLY1))) {
                finish_delivery<C>(dst_m, dst_e, first, lane0, back[kk], ez, acc[j][kk], eacc[j][kk]);
              } else {
                // the schedule keeps two deliveries to one child out of a
                // group, so this read-back of the adjoint written by an
                // earlier op of this group is never taken; kept for any schedule
                float acc_now[4];
                load4(dst_m, acc_now);
                finish_delivery<C>(dst_m, dst_e, false, lane0, back[kk], ez, acc_now, *dst_e);
              }
            }
          }
        }
      }
    }
  }
}

template <int C, int G, int TS>
__global__ void __launch_bounds__(TS *C, 3)
    k_adjoint_persistent(const NetDesc *__restrict__ nets, const WorkItem *__restrict__ work) {
  const WorkItem wi = work[blockIdx.x];
  const NetDesc &nd = nets[wi.net];
  adjoint_persistent_body<C, G, TS>(to_constant(nd.ops), to_constant(nd.group_begin), nd.n_groups, nd.L,
                                    to_global(nd.clv), to_global(nd.scl), to_global(nd.adj),
                                    to_global(nd.adj_exp), to_global(nd.tips), to_global(nd.Pf),
                                    to_global(nd.Qf), to_global(nd.tile_part), to_global(nd.freqs),
                                    to_global(nd.weights), to_global(nd.beta), nd.n_root, nd.n_uses,
                                    nd.n_pass, to_global(nd.grad_partials), wi.tile);
}

// ------------------------------------------- reverse sweep, two lanes per row
// The same sweep (ops, groups, flags, LDS pass slots, arithmetic) with each
// (site, category) row of 4 states held by two adjacent lanes: lane h = 0
// holds states 0, 1 and lane h = 1 states 2, 3 (threadIdx.x = (site C +
// category) 2 + h).  A tile then has twice the waves, each carrying half the
// registers per op, so more groups' loads are in flight per SIMD.  Products
// with a 4 x 4 matrix need the partner's two values (one DPP move each):
//   y = P x      my rows of P times (mine, partner's) x
//   back = P^T z partial sums over my two rows for all four columns; the
//                partner's partials for my columns arrive by DPP
//   Q y          as P x, on the exchanged y
// The P and Q tables in LDS are stored per lane half: block h of a (edge,
// category) holds rows 2h, 2h+1 with the columns reordered (2h, 2h+1, then
// the partner's two), so every lane reads its operands in one fixed order.
__device__ __forceinline__ float pair_swap(float v) { return dpp<DPP_XOR1>(v); }

// maximum / sum over the 2C lanes of a site (halves at lane bit 0,
// categories at bits 1 ..)
template <int C>
__device__ __forceinline__ float site_max2(float m) {  // integer max, as cat_maxf
  int u = __float_as_int(m);
  u = max(u, dpp_i<DPP_XOR1>(u));
  if constexpr (C > 1) u = max(u, dpp_i<DPP_XOR2>(u));
  if constexpr (C > 2) u = max(u, __shfl_xor(u, 4));
  if constexpr (C > 4) u = max(u, __shfl_xor(u, 8));
  return __int_as_float(u);
}
template <int C>
__device__ __forceinline__ double site_sum2(double m) {
  m += dpp<DPP_XOR1>(m);
  if constexpr (C > 1) m += dpp<DPP_XOR2>(m);
  if constexpr (C > 2) m += __shfl_xor(m, 4);
  if constexpr (C > 4) m += __shfl_xor(m, 8);
  return m;
}

// source index in a row-major 4 x 4 matrix of entry r (0..15) of lane half
// h's block: rows 2h + r / 4 % 2, columns (2h, 2h+1, partner's two)
__device__ __forceinline__ int half_block_src(int r) {
  const int h = r >> 3, k = (r >> 2) & 1, q = r & 3;
  const int row = 2 * h + k;
  const int col = q < 2 ? 2 * h + q : 2 * (1 - h) + (q - 2);
  return row * 4 + col;
}

// y (my two rows) = M (x, partner's x); M = the lane's 8-entry block
__device__ __forceinline__ void mv2(const float *__restrict__ M, float x0, float x1, float o0, float o1,
                                    float &y0, float &y1) {
  y0 = fmaf(M[3], o1, fmaf(M[2], o0, fmaf(M[1], x1, M[0] * x0)));
  y1 = fmaf(M[7], o1, fmaf(M[6], o0, fmaf(M[5], x1, M[4] * x0)));
}

// (P^T z) for my two states from my block and my two z values
__device__ __forceinline__ void mv2t(const float *__restrict__ M, float z0, float z1, float &b0, float &b1) {
  const float p0 = fmaf(M[4], z1, M[0] * z0);  // columns 2h, 2h+1: my states
  const float p1 = fmaf(M[5], z1, M[1] * z0);
  const float q0 = fmaf(M[6], z1, M[2] * z0);  // the partner's states
  const float q1 = fmaf(M[7], z1, M[3] * z0);
  b0 = p0 + pair_swap(q0);
  b1 = p1 + pair_swap(q1);
}

// x 2^ex + y 2^ey (two values), aligned to the larger exponent
__device__ __forceinline__ int align_add2(float &x0, float &x1, int ex, float y0, float y1, int ey) {
  const int e = max(ex, ey);
  const float sx = ldexpf(1.0f, ex - e), sy = ldexpf(1.0f, ey - e);
  x0 = fmaf(x0, sx, y0 * sy);
  x1 = fmaf(x1, sx, y1 * sy);
  return e;
}

__device__ __forceinline__ int renorm2(float &x0, float &x1, int ex, float m) {
  int e = 0;
  (void)frexpf(m, &e);
  x0 = ldexpf(x0, -e);
  x1 = ldexpf(x1, -e);
  return m > 0.f ? ex + e : EXP_EMPTY;
}

__device__ __forceinline__ void load2(const GAS float *__restrict__ p, float &a, float &b) {
  const float2 v = *reinterpret_cast<const GAS float2 *>(p);
  a = v.x;
  b = v.y;
}
__device__ __forceinline__ void store2(GAS float *__restrict__ p, float a, float b) {
  *reinterpret_cast<GAS float2 *>(p) = make_float2(a, b);
}

// delivered (b, eb) plus the accumulated (acc, ea) unless first, then
// renormalised (as finish_delivery)
template <int C>
__device__ __forceinline__ void finish_delivery2(float &b0, float &b1, int &eb, bool first, float a0, float a1,
                                                 int ea) {
  float mb = site_max2<C>(fmaxf(fabsf(b0), fabsf(b1)));
  if (!first) {
    eb = align_add2(b0, b1, mb > 0.f ? eb : EXP_EMPTY, a0, a1, ea);
    mb = site_max2<C>(fmaxf(fabsf(b0), fabsf(b1)));
  }
  eb = renorm2(b0, b1, eb, mb);
}

template <int C, int G, int TS>
__device__ __forceinline__ void adjoint_split_body(
    const CAS DevOp *__restrict__ ops, const CAS int32_t *__restrict__ group_begin, int n_groups,
    const DevLayout L, const GAS float *__restrict__ clv, const GAS int32_t *__restrict__ scl,
    GAS float *__restrict__ adj, GAS int32_t *__restrict__ adj_exp,
    const GAS uint8_t *__restrict__ tips, const GAS float *__restrict__ Pf,
    const GAS float *__restrict__ Qf, const GAS int32_t *__restrict__ tile_part,
    const GAS double *__restrict__ freqs, const GAS double *__restrict__ weights,
    const GAS double *__restrict__ beta, int n_root, int n_uses, int n_pass,
    GAS double *__restrict__ grad_partials, int tile) {
  float *sP = dyn_lds<float>();  // [edge][cat][half][8] then Q [cat][half][8], then pass slots
  constexpr int NT = TS * C * 2;
  constexpr int PWB = NT / 64;  // waves of this site block
  const int64_t site0 = static_cast<int64_t>(tile) * TS;
  const int part = tile_part[site0 / TILE];
  const int np = L.n_edges * C * 16;
  const GAS float *gP = Pf + static_cast<size_t>(part) * np;
  for (int i = threadIdx.x; i < np; i += NT) sP[i] = gP[(i & ~15) + half_block_src(i & 15)];
  for (int i = threadIdx.x; i < C * 16; i += NT) sP[np + i] = Qf[part * C * 16 + (i & ~15) + half_block_src(i & 15)];
  __syncthreads();
  const int h = threadIdx.x & 1;
  const int c = (threadIdx.x >> 1) % C;
  float Qr[8];  // this lane's half block of r_c Q, kept in registers
#pragma unroll
  for (int i = 0; i < 8; ++i) Qr[i] = sP[np + c * 16 + h * 8 + i];
  // pass slots: mantissas [slot][thread][2], site exponents [slot][site]
  float *pass_m = sP + ((np + C * 16 + 3) & ~3);
  int *pass_e = reinterpret_cast<int *>(pass_m + n_pass * NT * 2);
  const int site_in_tile = threadIdx.x / (2 * C);
  const int64_t site = site0 + site_in_tile;
  const int64_t n = L.n_sites_pad;
  const int wave = threadIdx.x / 64;
  const bool lead = (threadIdx.x & (2 * C - 1)) == 0;  // writes the site's exponent
  const double w = weights[site];
  const double pi0 = freqs[part * 4 + 2 * h], pi1 = freqs[part * 4 + 2 * h + 1];
  const size_t row = C * 4;
  const int off = c * 4 + 2 * h;  // my two states within a site's row
  for (int gi = n_groups - 1; gi >= 0; --gi) {
    const int g0 = group_begin[gi];
    const int nb = group_begin[gi + 1] - g0;
    float aq[G][2], x[G][2][2], xq[G][2], acc[G][2][2];
    int gq[G], kq[G], sk[G][2], eacc[G][2];
    uint32_t tm[G][2];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      if (j < nb) {
        const DevOp op = ops[g0 + j];
        kq[j] = scl[static_cast<size_t>(op.out) * n + site];
        if (op.flags & OP_HAS_CONSUMERS) {
          if (op.flags & OP_A_PASS) {
            const int slot = pass_slot(op.flags, 2);
            const float2 v = reinterpret_cast<const float2 *>(pass_m)[slot * NT + threadIdx.x];
            aq[j][0] = v.x;
            aq[j][1] = v.y;
            gq[j] = pass_e[slot * TS + site_in_tile];
          } else {
            load2(adj + (static_cast<size_t>(op.out) * n + site) * row + off, aq[j][0], aq[j][1]);
            gq[j] = adj_exp[static_cast<size_t>(op.out) * n + site];
          }
        }
        if (op.root >= 0) load2(clv + (static_cast<size_t>(op.out) * n + site) * row + off, xq[j][0], xq[j][1]);
      }
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        if (j < nb) {
          const DevOp op = ops[g0 + j];
          const int32_t kid = op.kid[kk];
          if (kk < op.nk && kid >= 0) {
            load2(clv + (static_cast<size_t>(kid) * n + site) * row + off, x[j][kk][0], x[j][kk][1]);
            sk[j][kk] = scl[static_cast<size_t>(kid) * n + site];
            if (op.flags & (kk == 0 ? OP_ACC_EARLY0 : OP_ACC_EARLY1)) {
              load2(adj + (static_cast<size_t>(kid) * n + site) * row + off, acc[j][kk][0], acc[j][kk][1]);
              eacc[j][kk] = adj_exp[static_cast<size_t>(kid) * n + site];
            }
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        if (j < nb) {
          const DevOp op = ops[g0 + j];
          const int32_t kid = op.kid[kk];
          if (kk < op.nk && kid < 0) tm[j][kk] = tips[static_cast<size_t>(-kid - 1) * n + site];
        }
      }
    }
    wait_loads();
#pragma unroll
    for (int j = 0; j < G; ++j) {
      if (j < nb) {
        const DevOp op = ops[g0 + j];
        int ks = 0;
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
          if (kk < op.nk) {
            if (op.kid[kk] < 0) {
              const uint32_t m = tm[j][kk] >> (2 * h);
              x[j][kk][0] = (m & 1u) ? 1.f : 0.f;
              x[j][kk][1] = (m & 2u) ? 1.f : 0.f;
            } else {
              ks += sk[j][kk];
            }
          }
        }
        const bool has_adj = (op.flags & OP_HAS_CONSUMERS) != 0;
        float a0 = has_adj ? aq[j][0] : 0.f, a1 = has_adj ? aq[j][1] : 0.f;
        int ea = has_adj ? gq[j] : EXP_EMPTY;
        if (op.root >= 0) {
          double t = site_sum2<C>(fma(pi0, static_cast<double>(xq[j][0]), pi1 * static_cast<double>(xq[j][1])));
          t *= 1.0 / C;
          const double b = beta[part * n_root + op.root] * w;
          int ef = EXP_EMPTY;
          float sd0 = 0.f, sd1 = 0.f;
          if (b != 0.0) {
            const double mf = frexp(b / (C * t), &ef);
            sd0 = static_cast<float>(mf * pi0);
            sd1 = static_cast<float>(mf * pi1);
          }
          ea = align_add2(a0, a1, ea, sd0, sd1, ef);
        }
        float y[2][2];
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
          if (kk < op.nk)
            mv2(sP + (op.edge[kk] * C + c) * 16 + h * 8, x[j][kk][0], x[j][kk][1], pair_swap(x[j][kk][0]),
                pair_swap(x[j][kk][1]), y[kk][0], y[kk][1]);
        const int ez = ea + kq[j] - ks;
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
          if (kk < op.nk) {
            const float z0 = op.nk == 2 ? a0 * y[1 - kk][0] : a0;
            const float z1 = op.nk == 2 ? a1 * y[1 - kk][1] : a1;
            float q0, q1, b0, b1;
            mv2(Qr, y[kk][0], y[kk][1], pair_swap(y[kk][0]), pair_swap(y[kk][1]), q0, q1);
            mv2t(sP + (op.edge[kk] * C + c) * 16 + h * 8, z0, z1, b0, b1);
            const float gd = fmaf(z1, q1, z0 * q0);
            const double gv = wave_sum_f32(ldexpf(gd, ez));
            if ((threadIdx.x & 63) == 0) grad_partials[grad_index(tile, n_uses, op.gslot[kk], PWB, wave)] = gv;
            const int32_t kid = op.kid[kk];
            if (kid < 0) continue;
            const bool first = (op.flags & (kk == 0 ? OP_FIRST_KID0 : OP_FIRST_KID1)) != 0;
            int eb = ez;
            if (op.flags & (kk == 0 ? OP_PASS_KID0 : OP_PASS_KID1)) {
              const int slot = pass_slot(op.flags, kk);
              float2 *sm = reinterpret_cast<float2 *>(pass_m) + slot * NT + threadIdx.x;
              int *se = pass_e + slot * TS + site_in_tile;
              float p0 = 0.f, p1 = 0.f;
              int pe = EXP_EMPTY;
              if (!first) {
                const float2 v = *sm;
                p0 = v.x;
                p1 = v.y;
                pe = *se;
              }
              finish_delivery2<C>(b0, b1, eb, first, p0, p1, pe);
              *sm = make_float2(b0, b1);
              *se = eb;  // every lane of the site writes the same value
            } else {
              GAS float *dst_m = adj + (static_cast<size_t>(kid) * n + site) * row + off;
              GAS int32_t *dst_e = adj_exp + static_cast<size_t>(kid) * n + site;
              float p0 = 0.f, p1 = 0.f;
              int pe = EXP_EMPTY;
              if (!first) {
                if (op.flags & (kk == 0 ? OP_ACC_EARLY0 : OP_ACC_EARLY1)) {
                  p0 = acc[j][kk][0];
                  p1 = acc[j][kk][1];
                  pe = eacc[j][kk];
                } else {
                  // an earlier op of this group or the previous group wrote it: read it now
                  load2(dst_m, p0, p1);
                  pe = *dst_e;
                }
              }
              finish_delivery2<C>(b0, b1, eb, first, p0, p1, pe);
              store2(dst_m, b0, b1);
              if (lead) *dst_e = eb;
            }
          }
        }
      }
    }
  }
}

template <int C, int G, int TS>
__global__ void __launch_bounds__(TS *C * 2)
    k_adjoint_split(const NetDesc *__restrict__ nets, const WorkItem *__restrict__ work) {
  const WorkItem wi = work[blockIdx.x];
  const NetDesc &nd = nets[wi.net];
  adjoint_split_body<C, G, TS>(to_constant(nd.ops), to_constant(nd.group_begin), nd.n_groups, nd.L,
                               to_global(nd.clv), to_global(nd.scl), to_global(nd.adj), to_global(nd.adj_exp),
                               to_global(nd.tips), to_global(nd.Pf), to_global(nd.Qf), to_global(nd.tile_part),
                               to_global(nd.freqs), to_global(nd.weights), to_global(nd.beta), nd.n_root,
                               nd.n_uses, nd.n_pass, to_global(nd.grad_partials), wi.tile);
}

// ============================================= FP32 form, 20 states (MFMA)
// CLV x P runs on the matrix cores.  For one rate category c and a tile of
// 32 sites, D = P_c X is a chain of 10 v_mfma_f32_32x32x2f32 (M = parent
// state, 20 padded to 32 with zero rows of P; K = child state, 2 per step;
// N = site).  Operand layout of that instruction, lane l, h = l / 32,
// col = l % 32:
//   A (32 x 2)   lane l holds A[col][h]
//   B (2 x 32)   lane l holds B[h][col]
//   D (32 x 32)  lane l, register r holds D[8 (r / 4) + 4 h + r % 4][col]
// CLVs and adjoints are stored per tile as [C][20 states][32 sites], so the
// B operand of step t (state 2t + h) and every D register (one state per
// half wave) is one coalesced pair of 128-byte rows; no transpose is
// needed between a CLV's producer and its consumers.  The A operands (P
// and P^T per partition, edge and category) are laid out in lane order by
// k_mfma_tables; r_c Q, the same for every edge, is read once per sweep.  A workgroup is C waves, wave c = rate
// category c of one 32-site tile.  CLV scalers are per site, as in the
// 4-state form: the waves combine their site maxima through LDS once per
// group of independent ops.  Adjoint exponents are per (site, category),
// so the reverse sweep only synchronises the waves at displayed-tree roots.
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int AA_STEPS = 10;   // K = 20 states, 2 per instruction
constexpr int AA_ROWS = 12;    // D registers holding real states (rows < 20 for h = 0)

__device__ __forceinline__ int aa_row(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }
__device__ __forceinline__ bool aa_valid(int r, int h) { return r < 8 || h == 0; }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void load_a(const GAS float *__restrict__ tab, int lane, float (&a)[AA_STEPS]) {
  const GAS float4 *p = reinterpret_cast<const GAS float4 *>(tab + lane * (AA_TABLE / 64));
  const float4 v0 = p[0], v1 = p[1], v2 = p[2];
  a[0] = v0.x; a[1] = v0.y; a[2] = v0.z; a[3] = v0.w;
  a[4] = v1.x; a[5] = v1.y; a[6] = v1.z; a[7] = v1.w;
  a[8] = v2.x; a[9] = v2.y;
}

// B operand of a child: rows 2t + h of column col ([20][32] block of one
// category), or the 0/1 vector of a tip's state mask
__device__ __forceinline__ void load_b(const GAS float *__restrict__ blk, int h, float (&b)[AA_STEPS]) {
#pragma unroll
  for (int t = 0; t < AA_STEPS; ++t) b[t] = blk[(2 * t + h) * AA_TILE];
}
__device__ __forceinline__ void tip_b(uint32_t mask, int h, float (&b)[AA_STEPS]) {
#pragma unroll
  for (int t = 0; t < AA_STEPS; ++t) b[t] = ((mask >> (2 * t + h)) & 1u) ? 1.0f : 0.0f;
}

__device__ __forceinline__ f32x16 mfma_chain(const float (&a)[AA_STEPS], const float (&b)[AA_STEPS]) {
  f32x16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < AA_STEPS; ++t) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[t], b[t], acc, 0, 0, 0);
  return acc;
}

// Operand tables: Pa[((p E + e) C + c) 3 + kind][lane][12], entry t < 10 of
// lane (h, col) = M[col][2t + h] with M = P (kind 0), P^T (1); zero
// for col >= 20 and t >= 10.
__global__ void k_mfma_tables(DevLayout L, const double *__restrict__ P, float *__restrict__ Pa) {
  const int64_t total = static_cast<int64_t>(L.n_parts) * L.n_edges * L.C * AA_KINDS * AA_TABLE;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int t = idx % 12;
  const int lane = (idx / 12) % 64;
  const int kind = (idx / AA_TABLE) % AA_KINDS;
  const int64_t mat = idx / (AA_TABLE * AA_KINDS);  // (p E + e) C + c
  const int h = lane >> 5, col = lane & 31;
  float v = 0.0f;
  if (t < AA_STEPS && col < AA_STATES) {
    const int k = 2 * t + h;
    const double *M = P + mat * (AA_STATES * AA_STATES);
    v = static_cast<float>(kind == 1 ? M[k * AA_STATES + col] : M[col * AA_STATES + k]);
  }
  Pa[idx] = v;
}

template <int C, int G>
__device__ __forceinline__ void forward_mfma_body(
    const CAS DevOp *__restrict__ ops, const CAS int32_t *__restrict__ group_begin, int n_groups,
    const DevLayout L, GAS float *__restrict__ clv, GAS int32_t *__restrict__ scl,
    const GAS uint32_t *__restrict__ tips, const GAS double *__restrict__ freqs,
    const GAS int32_t *__restrict__ tile_part, const GAS float *__restrict__ Pa_all,
    const GAS double *__restrict__ weights, GAS double *__restrict__ root_partials,
    int tile /* 32-site tile */) {
  __shared__ float smax[2][G][C][AA_TILE];
  __shared__ double ssum[2][G][C][AA_TILE];
  const int part = tile_part[tile >> 1];
  const int c = threadIdx.x >> 6, lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
  const int64_t n = L.n_sites_pad;
  const int64_t site = static_cast<int64_t>(tile) * AA_TILE + col;
  const size_t slot_stride = static_cast<size_t>(n) * C * AA_STATES;
  const size_t tile_off = static_cast<size_t>(tile) * (C * AA_STATES * AA_TILE) +
                          static_cast<size_t>(c) * (AA_STATES * AA_TILE) + col;
  const GAS float *__restrict__ Pa =
      Pa_all + static_cast<size_t>(part) * L.n_edges * C * AA_KINDS * AA_TABLE;
  const double w = weights[site];
  const size_t n_wt = static_cast<size_t>(L.n_tiles) * 2 * C;
  const size_t wt = static_cast<size_t>(tile) * C + c;
  for (int gi = 0; gi < n_groups; ++gi) {
    const int buf = gi & 1;
    const int g0 = group_begin[gi];
    const int nb = group_begin[gi + 1] - g0;
    float v[G][AA_ROWS];
    int ks[G];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      ks[j] = 0;
      if (j < nb) {
        const DevOp op = ops[g0 + j];
#pragma unroll
        for (int r = 0; r < AA_ROWS; ++r) v[j][r] = 1.0f;
        // both children's operands as raw values first, then the chains
        float a[2][AA_STEPS], b[2][AA_STEPS];
        int sk[2];
        uint32_t tm[2];
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
          // the tip mask and the scaler are loaded for every child (tip 0
          // and variant 0 stand in for the absent one): a value loaded on
          // one path only would be copied at the join, which waits for
          // every load issued so far
          const int32_t kid = kk < op.nk ? op.kid[kk] : -1;
          tm[kk] = tips[static_cast<size_t>(kid < 0 ? -kid - 1 : 0) * n + site];
          sk[kk] = scl[static_cast<size_t>(kid < 0 ? 0 : kid) * n + site];
          if (kk < op.nk) {
            load_a(Pa + (static_cast<size_t>(op.edge[kk]) * C + c) * AA_KINDS * AA_TABLE, lane, a[kk]);
            if (kid >= 0) load_b(clv + kid * slot_stride + tile_off, h, b[kk]);
          }
        }
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
          if (kk < op.nk) {
            if (op.kid[kk] < 0) tip_b(tm[kk], h, b[kk]);
            ks[j] += op.kid[kk] < 0 ? 0 : sk[kk];
            const f32x16 d = mfma_chain(a[kk], b[kk]);
#pragma unroll
            for (int r = 0; r < AA_ROWS; ++r) v[j][r] *= d[r];
          }
        }
        float m = 0.0f;
#pragma unroll
        for (int r = 0; r < AA_ROWS; ++r) m = fmaxf(m, v[j][r]);
        m = fmaxf(m, __shfl_xor(m, 32));
        if (h == 0) smax[buf][j][c][col] = m;
        if (op.root >= 0) {
          double s = 0.0;
#pragma unroll
          for (int r = 0; r < AA_ROWS; ++r)
            if (aa_valid(r, h)) s = fma(freqs[part * AA_STATES + aa_row(r, h)], static_cast<double>(v[j][r]), s);
          s += __shfl_xor(s, 32);
          if (h == 0) ssum[buf][j][c][col] = s;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < G; ++j) {
      if (j < nb) {
        const DevOp op = ops[g0 + j];
        float m = 0.0f;
#pragma unroll
        for (int cc = 0; cc < C; ++cc) m = fmaxf(m, smax[buf][j][cc][col]);
        int e = 0;
        if (m > 0.f) (void)frexpf(m, &e);
        const int k = ks[j] - e;
        GAS float *dst = clv + op.out * slot_stride + tile_off;
#pragma unroll
        for (int r = 0; r < AA_ROWS; ++r)
          if (aa_valid(r, h)) dst[aa_row(r, h) * AA_TILE] = ldexpf(v[j][r], -e);
        if (c == 0 && h == 0) scl[static_cast<size_t>(op.out) * n + site] = k;
        if (op.root >= 0) {
          double t = 0.0;
#pragma unroll
          for (int cc = 0; cc < C; ++cc) t += ssum[buf][j][cc][col];
          t = ldexp(t, -e) * (1.0 / C);
          double val = (c == 0 && h == 0 && w != 0.0) ? w * (log(t) - k * LN2) : 0.0;
          val = wave_sum(val);
          if (lane == 0) root_partials[static_cast<size_t>(op.root) * n_wt + wt] = val;
        }
      }
    }
  }
}

template <int C, int G>
__global__ void __launch_bounds__(64 * C)
    k_forward_mfma(const NetDesc *__restrict__ nets, const WorkItem *__restrict__ work) {
  const WorkItem wi = work[blockIdx.x];
  const NetDesc &nd = nets[wi.net];
  forward_mfma_body<C, G>(to_constant(nd.ops), to_constant(nd.group_begin), nd.n_groups, nd.L,
                          to_global(nd.clv), to_global(nd.scl),
                          to_global(reinterpret_cast<const uint32_t *>(nd.tips)), to_global(nd.freqs),
                          to_global(nd.tile_part), to_global(nd.Pa), to_global(nd.weights),
                          to_global(nd.root_partials), wi.tile);
}

// Adjoint delivery to a child (push form) with one exponent per (site,
// category): lanes l and l ^ 32 hold the two halves of a site's states.
__device__ __forceinline__ void deliver_adjoint_mfma(GAS float *__restrict__ dst_m, GAS int32_t *__restrict__ dst_e,
                                                     bool first, int h, float (&b)[AA_ROWS], int eb) {
  float mb = 0.0f;
#pragma unroll
  for (int r = 0; r < AA_ROWS; ++r) mb = fmaxf(mb, fabsf(b[r]));
  mb = fmaxf(mb, __shfl_xor(mb, 32));
  if (!(mb > 0.f)) eb = EXP_EMPTY;
  if (!first) {
    float acc[AA_ROWS];
#pragma unroll
    for (int r = 0; r < AA_ROWS; ++r) acc[r] = aa_valid(r, h) ? dst_m[aa_row(r, h) * AA_TILE] : 0.0f;
    const int ea = *dst_e;
    if (ea > eb) {
#pragma unroll
      for (int r = 0; r < AA_ROWS; ++r) b[r] = acc[r] + ldexpf(b[r], eb - ea);
      eb = ea;
    } else {
#pragma unroll
      for (int r = 0; r < AA_ROWS; ++r) b[r] += ldexpf(acc[r], ea - eb);
    }
    mb = 0.0f;
#pragma unroll
    for (int r = 0; r < AA_ROWS; ++r) mb = fmaxf(mb, fabsf(b[r]));
    mb = fmaxf(mb, __shfl_xor(mb, 32));
  }
  if (mb > 0.f) {
    int e;
    (void)frexpf(mb, &e);
#pragma unroll
    for (int r = 0; r < AA_ROWS; ++r) b[r] = ldexpf(b[r], -e);
    eb += e;
  } else {
    eb = EXP_EMPTY;
  }
#pragma unroll
  for (int r = 0; r < AA_ROWS; ++r)
    if (aa_valid(r, h)) dst_m[aa_row(r, h) * AA_TILE] = b[r];
  if (h == 0) *dst_e = eb;
}

// Reverse sweep of the 20-state FP32 form (push form, as k_adjoint_persistent):
// y_k = P_k x_k from the child's B operand; P'_k x_k = r_c Q y_k and
// P_k^T z_k (z_k = Abar * y_other) take y_k and z_k from the D layout into
// B operands through wave-private LDS blocks, and P_k^T z_k is delivered to
// the child.
template <int C>
__device__ __forceinline__ void adjoint_mfma_body(
    const CAS DevOp *__restrict__ ops, const CAS int32_t *__restrict__ group_begin, int n_groups,
    const DevLayout L, const GAS float *__restrict__ clv, const GAS int32_t *__restrict__ scl,
    GAS float *__restrict__ adj, GAS int32_t *__restrict__ adj_exp,
    const GAS uint32_t *__restrict__ tips, const GAS double *__restrict__ freqs,
    const GAS int32_t *__restrict__ tile_part, const GAS float *__restrict__ Pa_all,
    const GAS float *__restrict__ Qf, const GAS double *__restrict__ weights, const GAS double *__restrict__ beta,
    int n_root, int n_uses, GAS double *__restrict__ grad_partials, int tile) {
  __shared__ double rsum[2][C][AA_TILE];
  __shared__ float zb[C][2][AA_STATES * AA_TILE];
  const int part = tile_part[tile >> 1];
  const int c = threadIdx.x >> 6, lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
  const int64_t n = L.n_sites_pad;
  const int64_t site = static_cast<int64_t>(tile) * AA_TILE + col;
  const size_t slot_stride = static_cast<size_t>(n) * C * AA_STATES;
  const size_t tile_off = static_cast<size_t>(tile) * (C * AA_STATES * AA_TILE) +
                          static_cast<size_t>(c) * (AA_STATES * AA_TILE) + col;
  // adjoint exponents: [slot][tile][C][32]
  const size_t exp_stride = static_cast<size_t>(n) * C;
  const size_t exp_off = static_cast<size_t>(tile) * (C * AA_TILE) + static_cast<size_t>(c) * AA_TILE + col;
  const GAS float *__restrict__ Pa =
      Pa_all + static_cast<size_t>(part) * L.n_edges * C * AA_KINDS * AA_TABLE;
  const double w = weights[site];
  float *zw = zb[c][0], *yw = zb[c][1];
  // A operand of r_c Q (the partition's and this wave's category), held for
  // the whole sweep: P' x = r_c Q (P x) reuses y = P x instead of a P' table
  float qa[AA_STEPS];
  {
    const GAS float *q = Qf + (static_cast<size_t>(part) * C + c) * (AA_STATES * AA_STATES);
#pragma unroll
    for (int t = 0; t < AA_STEPS; ++t) qa[t] = col < AA_STATES ? q[col * AA_STATES + 2 * t + h] : 0.0f;
  }
  int rb = 0;
  for (int gi = n_groups - 1; gi >= 0; --gi) {
    const int g0 = group_begin[gi];
    const int g1 = group_begin[gi + 1];
    for (int oi = g0; oi < g1; ++oi) {
      const DevOp op = ops[oi];
      const int kq = scl[static_cast<size_t>(op.out) * n + site];
      float a[AA_ROWS];
      int ea = EXP_EMPTY;
      if (op.flags & OP_HAS_CONSUMERS) {
        const GAS float *src = adj + op.out * slot_stride + tile_off;
#pragma unroll
        for (int r = 0; r < AA_ROWS; ++r) a[r] = aa_valid(r, h) ? src[aa_row(r, h) * AA_TILE] : 0.0f;
        ea = adj_exp[op.out * exp_stride + exp_off];
      } else {
#pragma unroll
        for (int r = 0; r < AA_ROWS; ++r) a[r] = 0.0f;
      }
      float b[2][AA_STEPS];
      int skraw[2];
      uint32_t tm[2];
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        // the tip mask and the scaler are loaded for every child (tip 0 and
        // variant 0 stand in for the absent one): a value loaded on one
        // path only would be copied at the join, which waits for every
        // load issued so far
        const int32_t kid = kk < op.nk ? op.kid[kk] : -1;
        tm[kk] = tips[static_cast<size_t>(kid < 0 ? -kid - 1 : 0) * n + site];
        skraw[kk] = scl[static_cast<size_t>(kid < 0 ? 0 : kid) * n + site];
        if (kk < op.nk && kid >= 0) load_b(clv + kid * slot_stride + tile_off, h, b[kk]);
      }
      // every A operand of the op (P and, for inner children, P^T) is
      // requested here with the op's other loads, so the chains below wait
      // for one round trip per op instead of one per table
      float pa[2][AA_STEPS], pt[2][AA_STEPS];
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        if (kk < op.nk) {
          const GAS float *tab = Pa + (static_cast<size_t>(op.edge[kk]) * C + c) * AA_KINDS * AA_TABLE;
          load_a(tab, lane, pa[kk]);
          if (op.kid[kk] >= 0) load_a(tab + AA_TABLE, lane, pt[kk]);
        }
      }
      int sk[2] = {0, 0};
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        if (kk < op.nk) {
          if (op.kid[kk] < 0)
            tip_b(tm[kk], h, b[kk]);
          else
            sk[kk] = skraw[kk];
        }
      }
      const int ks = sk[0] + sk[1];
      if (op.root >= 0) {
        // seed beta * w * (1/C) pi_i / Ltilde, Ltilde = (1/C) sum_c pi . x_c
        const GAS float *xq = clv + op.out * slot_stride + tile_off;
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < AA_ROWS; ++r)
          if (aa_valid(r, h))
            s = fma(freqs[part * AA_STATES + aa_row(r, h)], static_cast<double>(xq[aa_row(r, h) * AA_TILE]), s);
        s += __shfl_xor(s, 32);
        if (h == 0) rsum[rb][c][col] = s;
        __syncthreads();
        double t = 0.0;
#pragma unroll
        for (int cc = 0; cc < C; ++cc) t += rsum[rb][cc][col];
        t *= 1.0 / C;
        rb ^= 1;
        const double bw = beta[part * n_root + op.root] * w;
        int ef = EXP_EMPTY;
        float sd[AA_ROWS];
        double mf = 0.0;
        if (bw != 0.0) mf = frexp(bw / (C * t), &ef);
#pragma unroll
        for (int r = 0; r < AA_ROWS; ++r)
          sd[r] = aa_valid(r, h) ? static_cast<float>(mf * freqs[part * AA_STATES + aa_row(r, h)]) : 0.0f;
        if (ef > ea) {
#pragma unroll
          for (int r = 0; r < AA_ROWS; ++r) a[r] = sd[r] + ldexpf(a[r], ea - ef);
          ea = ef;
        } else {
#pragma unroll
          for (int r = 0; r < AA_ROWS; ++r) a[r] += ldexpf(sd[r], ef - ea);
        }
      }
      f32x16 y[2];
#pragma unroll
      for (int kk = 0; kk < 2; ++kk)
        if (kk < op.nk) y[kk] = mfma_chain(pa[kk], b[kk]);
      const int ez = ea + kq - ks;
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        if (kk < op.nk) {
          float z[AA_ROWS];
#pragma unroll
          for (int r = 0; r < AA_ROWS; ++r) z[r] = op.nk == 2 ? a[r] * y[1 - kk][r] : a[r];
          // y_k from the D layout into a B operand (rows 2t + h) through the
          // wave's LDS block, then P'_k x_k = r_c Q y_k
#pragma unroll
          for (int r = 0; r < AA_ROWS; ++r)
            if (aa_valid(r, h)) yw[aa_row(r, h) * AA_TILE + col] = y[kk][r];
          wave_sync();
          float yb[AA_STEPS];
#pragma unroll
          for (int t = 0; t < AA_STEPS; ++t) yb[t] = yw[(2 * t + h) * AA_TILE + col];
          wave_sync();
          const f32x16 dy = mfma_chain(qa, yb);
          float gd = 0.0f;
#pragma unroll
          for (int r = 0; r < AA_ROWS; ++r) gd = fmaf(z[r], dy[r], gd);
          const double gv = wave_sum(ldexp(static_cast<double>(gd), ez));
          if (lane == 0) grad_partials[grad_index(tile, n_uses, op.gslot[kk], C, c)] = gv;
          const int32_t kid = op.kid[kk];
          if (kid >= 0) {
#pragma unroll
            for (int r = 0; r < AA_ROWS; ++r)
              if (aa_valid(r, h)) zw[aa_row(r, h) * AA_TILE + col] = z[r];
            wave_sync();
            float bz[AA_STEPS];
#pragma unroll
            for (int t = 0; t < AA_STEPS; ++t) bz[t] = zw[(2 * t + h) * AA_TILE + col];
            wave_sync();
            const f32x16 bk = mfma_chain(pt[kk], bz);
            float back[AA_ROWS];
#pragma unroll
            for (int r = 0; r < AA_ROWS; ++r) back[r] = bk[r];
            deliver_adjoint_mfma(adj + kid * slot_stride + tile_off, adj_exp + kid * exp_stride + exp_off,
                                 (op.flags & (kk == 0 ? OP_FIRST_KID0 : OP_FIRST_KID1)) != 0, h, back, ez);
          }
        }
      }
    }
  }
}

template <int C>
__global__ void __launch_bounds__(64 * C)
    k_adjoint_mfma(const NetDesc *__restrict__ nets, const WorkItem *__restrict__ work) {
  const WorkItem wi = work[blockIdx.x];
  const NetDesc &nd = nets[wi.net];
  adjoint_mfma_body<C>(to_constant(nd.ops), to_constant(nd.group_begin), nd.n_groups, nd.L,
                       to_global(nd.clv), to_global(nd.scl), to_global(nd.adj), to_global(nd.adj_exp),
                       to_global(reinterpret_cast<const uint32_t *>(nd.tips)), to_global(nd.freqs),
                       to_global(nd.tile_part), to_global(nd.Pa), to_global(nd.Qf), to_global(nd.weights),
                       to_global(nd.beta), nd.n_root, nd.n_uses, to_global(nd.grad_partials), wi.tile);
}

// ===================================================== branch derivatives
// Sumtables of one branch e, one row per displayed tree t containing it
// (the device-side counterpart of computePartitionSumtables,
// src/likelihood/LikelihoodDerivatives.cpp:249-344).  A path adjoint walks
// from t's root CLV down to e with the seed (1/C) pi_i / Ltilde (weight 1),
// so at e the vector z = d log L_s / d (P_e x) satisfies z . (P_e x) = 1.
// In the eigenbasis of the partition's rate matrix
//   ST[s,c,k] = (z^T V)_k (V^-1 x)_k,   rho_s(b) = sum_{c,k} ST e^(lambda_k r_c b)
// and rho_s(b) = L_s(b) / L_s(b_now).  Works on either form's CLVs (T =
// double with 2^256 scaler units, or float with unit exponents) in FP64.
template <int S>
__device__ __forceinline__ void load_any(const double *p, double (&x)[S]) {
  load_clv<S>(p, x);
}
template <int S>
__device__ __forceinline__ void load_any(const float *p, double (&x)[S]) {
#pragma unroll
  for (int i = 0; i < S; ++i) x[i] = static_cast<double>(p[i]);
}

// The P blocks (child edge, sibling edge; every category) of a path's first
// n_staged steps into LDS, [step][2][C][S][S]; ends with a barrier.
template <int S, int C>
__device__ __forceinline__ void stage_path_p(double *sPp, const PathStep *__restrict__ steps, int n_staged,
                                             const double *__restrict__ P, int part, const DevLayout &L) {
  if constexpr (S == 4) {
    constexpr int ROW = C * S * S;
    for (int i = threadIdx.x; i < n_staged * 2 * ROW; i += blockDim.x) {
      const int blk = i / ROW, r = i % ROW;
      const PathStep &ps = steps[blk >> 1];
      const int edge = (blk & 1) ? ps.edge_sib : ps.edge_child;
      sPp[i] = P[(static_cast<size_t>(part) * L.n_edges + edge) * ROW + r];
    }
    __syncthreads();
  }
}

// Sumtable row of a path at (site, category): the path adjoint z walked
// from the displayed tree's root to the branch, then (z^T V)_k (V^-1 x)_k
// with x the CLV below the branch.  The steps go in batches of PB: every
// load of a batch (sibling rows, tip masks, scalers; slot 0 and tip 0 stand
// in for absent ones, so no load sits on one path of a branch) is issued
// before its arithmetic, which then only waits on LDS and registers.  A
// step's own scaler is the previous step's child scaler.
template <int S, int C, typename T>
__device__ __forceinline__ void path_sumtable(const PathStep *__restrict__ steps, int b0, int b1,
                                              const double *sPp, int n_staged, const DevLayout &L,
                                              int part, int64_t site, int c, const T *__restrict__ clv,
                                              const int32_t *__restrict__ scl, int scale_unit,
                                              const typename TipType<S>::type *__restrict__ tips,
                                              const double *__restrict__ P, const double *__restrict__ V,
                                              const double *__restrict__ Vinv,
                                              const double *__restrict__ freqs, double (&st)[S]) {
  using TT = typename TipType<S>::type;
  constexpr int PB = S == 4 ? 4 : 1;
  const int64_t n = L.n_sites_pad;
  const size_t row = static_cast<size_t>(C * S);
  auto slot_of = [](int32_t x) { return x >= 0 ? x : 0; };
  auto tip_of = [](int32_t x) { return (x < 0 && x != NO_CHILD) ? -x - 1 : 0; };
  auto prow = [&](int j, int which, int edge) -> const double * {
    if (S == 4 && j < n_staged) return sPp + ((j * 2 + which) * C + c) * (S * S);
    return P + ((static_cast<size_t>(part) * L.n_edges + edge) * C + c) * (S * S);
  };
  const int32_t root = steps[b0].out, low = steps[b1 - 1].child;
  double xr[S], xc[S];
  load_any<S>(clv + (static_cast<size_t>(root) * n + site) * row + c * S, xr);
  int kq = scl[static_cast<size_t>(root) * n + site];
  load_any<S>(clv + (static_cast<size_t>(slot_of(low)) * n + site) * row + c * S, xc);
  const TT tmc = tips[static_cast<size_t>(tip_of(low)) * n + site];
  double a[S];
  {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < S; ++i) t = fma(freqs[part * S + i], xr[i], t);
    t = cat_sum<C>(t) * (1.0 / C);
#pragma unroll
    for (int i = 0; i < S; ++i) a[i] = freqs[part * S + i] / (C * t);
  }
  for (int s0 = b0; s0 < b1; s0 += PB) {
    double xs[PB][S];
    int kc[PB], ks[PB];
    TT tm[PB];
#pragma unroll
    for (int u = 0; u < PB; ++u) {
      if (s0 + u < b1) {
        const PathStep &ps = steps[s0 + u];
        kc[u] = scl[static_cast<size_t>(slot_of(ps.child)) * n + site];
        ks[u] = scl[static_cast<size_t>(slot_of(ps.sib)) * n + site];
        tm[u] = tips[static_cast<size_t>(tip_of(ps.sib)) * n + site];
        load_any<S>(clv + (static_cast<size_t>(slot_of(ps.sib)) * n + site) * row + c * S, xs[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < PB; ++u) {
      if (s0 + u < b1) {
        const PathStep &ps = steps[s0 + u];
        const int j = s0 + u - b0;
        double y[S];
        int ksu = 0;
        if (ps.nk == 2) {
          if (ps.sib < 0)
            tip_vec<S>(tm[u], xs[u]);
          else
            ksu = ks[u];
          matvec<S>(prow(j, 1, ps.edge_sib), xs[u], y);
        } else {
#pragma unroll
          for (int i = 0; i < S; ++i) y[i] = 1.0;
        }
        const int kcu = ps.child >= 0 ? kc[u] : 0;
        const double F = ldexp(1.0, scale_unit * (kq - kcu - ksu));
        double z[S];
#pragma unroll
        for (int i = 0; i < S; ++i) z[i] = a[i] * F * y[i];
        if (s0 + u + 1 < b1) {
          matvec_t<S>(prow(j, 0, ps.edge_child), z, a);  // adjoint of the on-path child
          kq = kcu;
        } else {
          double uu[S], v[S];
          if (low < 0) tip_vec<S>(tmc, xc);
          matvec_t<S>(V + static_cast<size_t>(part) * S * S, z, uu);  // u_k = sum_i z_i V[i][k]
          matvec<S>(Vinv + static_cast<size_t>(part) * S * S, xc, v);  // v_k = sum_j Vinv[k][j] x_j
#pragma unroll
          for (int kk = 0; kk < S; ++kk) st[kk] = uu[kk] * v[kk];
        }
      }
    }
  }
}

template <int S, int C, typename T>
__global__ void __launch_bounds__(TILE *C)
    k_sumtables(const PathStep *__restrict__ steps, const int32_t *__restrict__ path_begin,
                int n_paths, DevLayout L, const int32_t *__restrict__ tile_part, const T *__restrict__ clv,
                const int32_t *__restrict__ scl, int scale_unit,
                const typename TipType<S>::type *__restrict__ tips,
                const double *__restrict__ P, const double *__restrict__ V,
                const double *__restrict__ Vinv, const double *__restrict__ freqs,
                double *__restrict__ st) {
  double *sPp = dyn_lds<double>();
  const int tile = blockIdx.x;
  const int part = tile_part[tile];
  const int c = threadIdx.x % C;
  const int64_t site = static_cast<int64_t>(tile) * TILE + threadIdx.x / C;
  const size_t row = static_cast<size_t>(C * S);
  // one displayed tree's path per workgroup (grid tiles x paths): the path
  // walks are independent, and one per workgroup fills the GPU
  const int j = blockIdx.y;
  if (j >= n_paths) return;
  const int b0 = path_begin[j], b1 = path_begin[j + 1];
  const int n_staged = S == 4 ? min(b1 - b0, PATH_STAGED_STEPS) : 0;
  stage_path_p<S, C>(sPp, steps + b0, n_staged, P, part, L);
  double stv[S];
  path_sumtable<S, C, T>(steps, b0, b1, sPp, n_staged, L, part, site, c, clv, scl, scale_unit, tips, P, V, Vinv,
                         freqs, stv);
  double *out = st + (static_cast<size_t>(j) * L.n_sites_pad + site) * row + c * S;
#pragma unroll
  for (int kk = 0; kk < S; ++kk) out[kk] = stv[kk];
}

// Per displayed tree j and wave: sums over sites of w log rho(b),
// w rho'/rho and w (rho''/rho - (rho'/rho)^2) at branch length b, from the
// sumtable row of (site, category).
template <int S, int C>
__device__ __forceinline__ void sumtable_partials(const double (&in)[S], const double (&ex)[S],
                                                  const double (&lr)[S], double w, int c, double *out3) {
  double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const double q = in[k] * ex[k];
    t0 += q;
    t1 += q * lr[k];
    t2 += q * lr[k] * lr[k];
  }
  t0 = cat_sum<C>(t0);
  t1 = cat_sum<C>(t1);
  t2 = cat_sum<C>(t2);
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  if (c == 0 && w != 0.0) {
    const double r1 = t1 / t0;
    v0 = w * log(t0);
    v1 = w * r1;
    v2 = w * (t2 / t0 - r1 * r1);
  }
  out3[0] = wave_sum(v0);
  out3[1] = wave_sum(v1);
  out3[2] = wave_sum(v2);
}

// lambda_k r_c and e^(lambda_k r_c b) of the workgroup's partition for every
// (category, state): site-independent, so computed once per workgroup (the
// exponentials would otherwise dominate the evaluation's arithmetic) and
// handed to every thread through LDS; the values are the ones each thread
// computed itself before.  Ends with a barrier.
template <int S, int C>
__device__ __forceinline__ void stage_exps(const double *__restrict__ ev, const double *__restrict__ rates, int part,
                                           double brlen, int c, double (&ex)[S], double (&lr)[S]) {
  __shared__ double s_lr[C * S], s_ex[C * S];
  for (int i = threadIdx.x; i < C * S; i += blockDim.x) {
    const double l = ev[part * S + i % S] * rates[part * C + i / S];
    s_lr[i] = l;
    s_ex[i] = exp(l * brlen);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < S; ++k) {
    lr[k] = s_lr[c * S + k];
    ex[k] = s_ex[c * S + k];
  }
}

template <int S, int C>
__global__ void __launch_bounds__(TILE *C)
    k_sumtable_eval(const double *__restrict__ st, int n_paths, DevLayout L,
                    const int32_t *__restrict__ tile_part,
                    const double *__restrict__ ev, const double *__restrict__ rates,
                    const double *__restrict__ weights, double brlen,
                    double *__restrict__ partials) {
  if (static_cast<int>(blockIdx.y) >= n_paths) return;
  constexpr int PW = TILE * C / 64;
  const int tile = blockIdx.x;
  const int part = tile_part[tile];
  const int c = threadIdx.x % C;
  const int64_t site = static_cast<int64_t>(tile) * TILE + threadIdx.x / C;
  const int64_t n = L.n_sites_pad;
  const size_t row = static_cast<size_t>(C * S);
  const size_t n_wt = static_cast<size_t>(L.n_tiles) * PW;
  const int wt = tile * PW + threadIdx.x / 64;
  const double w = weights[site];
  double ex[S], lr[S];
  stage_exps<S, C>(ev, rates, part, brlen, c, ex, lr);
  // one path per workgroup (grid tiles x paths)
  for (int j = blockIdx.y; j < min(n_paths, static_cast<int>(blockIdx.y) + 1); ++j) {
    const double *in = st + (static_cast<size_t>(j) * n + site) * row + c * S;
    double stv[S], v[3];
#pragma unroll
    for (int k = 0; k < S; ++k) stv[k] = in[k];
    sumtable_partials<S, C>(stv, ex, lr, w, c, v);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
      for (int q = 0; q < 3; ++q) partials[(static_cast<size_t>(j) * 3 + q) * n_wt + wt] = v[q];
  }
}

// ------------------------------------- Newton-Raphson of a branch on the device
// The sumtable evaluation and its reduction at the length held in the
// device state (nothing once the iteration is done, unless `final_eval`),
// the Newton step and the acceptance check.  The step restates the
// safeguarded iteration of the host sweep (netrax_amd/optimization.py
// newton_maximise) in FP64; the mixture of the per-tree derivatives is
// computePartitionLhData's (LikelihoodDerivatives.cpp:30-188, reference
// normalisation) or the true derivative.
template <int S, int C>
__global__ void __launch_bounds__(TILE *C)
    k_sumtable_eval_nr(const double *__restrict__ st, DevLayout L, const int32_t *__restrict__ tile_part,
                       const double *__restrict__ ev, const double *__restrict__ rates,
                       const double *__restrict__ weights, const NewtonState *__restrict__ ns, int final_eval,
                       double *__restrict__ partials) {
  if (!final_eval && ns->done) return;
  constexpr int PW = TILE * C / 64;
  const int n_paths = ns->n_paths;
  if (static_cast<int>(blockIdx.y) >= n_paths) return;  // uniform: the grid covers the largest path count
  const double brlen = ns->x;
  const int tile = blockIdx.x;
  const int part = tile_part[tile];
  const int c = threadIdx.x % C;
  const int64_t site = static_cast<int64_t>(tile) * TILE + threadIdx.x / C;
  const int64_t n = L.n_sites_pad;
  const size_t row = static_cast<size_t>(C * S);
  const size_t n_wt = static_cast<size_t>(L.n_tiles) * PW;
  const int wt = tile * PW + threadIdx.x / 64;
  const double w = weights[site];
  double ex[S], lr[S];
  stage_exps<S, C>(ev, rates, part, brlen, c, ex, lr);
  // one path per workgroup (grid tiles x the largest path count)
  for (int j = blockIdx.y; j < min(n_paths, static_cast<int>(blockIdx.y) + 1); ++j) {
    const double *in = st + (static_cast<size_t>(j) * n + site) * row + c * S;
    double stv[S], v[3];
#pragma unroll
    for (int k = 0; k < S; ++k) stv[k] = in[k];
    sumtable_partials<S, C>(stv, ex, lr, w, c, v);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
      for (int q = 0; q < 3; ++q) partials[(static_cast<size_t>(j) * 3 + q) * n_wt + wt] = v[q];
  }
}

__global__ void __launch_bounds__(256)
    k_reduce_segments_nr(const double *__restrict__ in, size_t row_stride, int pw,
                         const int32_t *__restrict__ part_tile_begin, int P, const NewtonState *__restrict__ ns,
                         int final_eval, double *__restrict__ out) {
  if (!final_eval && ns->done) return;
  if (static_cast<int>(blockIdx.x) >= ns->n_paths * 3 * P) return;
  __shared__ double sh[4];
  const int row = blockIdx.x / P, p = blockIdx.x % P;
  const double *r = in + static_cast<size_t>(row) * row_stride;
  const int t0 = part_tile_begin[p] * pw, t1 = part_tile_begin[p + 1] * pw;
  double s = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += 256) s += r[t];
  s = block_sum<256>(s, sh);
  if (threadIdx.x == 0) out[static_cast<size_t>(row) * P + p] = s;
}

// tree lnL of class k in partition p at the evaluated length
__device__ __forceinline__ double nr_tree_logl(const double *sums, const DerivClassDev &cl, const double *base,
                                               int P, int nc, int k, int p) {
  double ll = base[static_cast<size_t>(p) * nc + k];
  if (cl.has && cl.path >= 0) ll += sums[(static_cast<size_t>(cl.path) * 3 + 0) * P + p];
  return ll;
}

// The class table, the base tree lnLs and the path sums are small: the
// workgroup copies them into LDS when they fit (NR_LDS_BYTES), so the serial
// mixture below reads LDS instead of making a memory round trip per value.
constexpr int NR_LDS_BYTES = 32768;
__device__ __forceinline__ void nr_stage(const double *&sums, const DerivClassDev *&cls, const double *&base, int P,
                                         int nc, int n_paths) {
  const size_t n_cls = static_cast<size_t>(nc), n_base = static_cast<size_t>(P) * nc,
               n_sums = static_cast<size_t>(n_paths) * 3 * P;
  if (sizeof(DerivClassDev) * n_cls + sizeof(double) * (n_base + n_sums) > NR_LDS_BYTES) return;
  DerivClassDev *s_cls = dyn_lds<DerivClassDev>();
  double *s_base = reinterpret_cast<double *>(s_cls + n_cls), *s_sums = s_base + n_base;
  for (size_t i = threadIdx.x; i < n_cls; i += blockDim.x) s_cls[i] = cls[i];
  for (size_t i = threadIdx.x; i < n_base; i += blockDim.x) s_base[i] = base[i];
  for (size_t i = threadIdx.x; i < n_sums; i += blockDim.x) s_sums[i] = sums[i];
  __syncthreads();
  cls = s_cls;
  base = s_base;
  sums = s_sums;
}

// Butterfly reductions over the 64 lanes of the (one-wave) Newton kernels;
// every lane ends with the result.
__device__ __forceinline__ double nr_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double nr_wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int nr_wave_isum(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ int nr_wave_imax(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// One wave: lane l takes the classes l, l + 64, ...; the class loops of the
// mixture become wave reductions.
__global__ void __launch_bounds__(64)
    k_newton_step(const double *__restrict__ sums_g, int P, int nc,
                  const DerivClassDev *__restrict__ cls_g, const double *__restrict__ base_g,
                  int variant, int mode, NewtonState *__restrict__ ns) {
  if (ns->done) return;  // uniform: every thread reads the same flag
  const double *sums = sums_g, *base = base_g;
  const DerivClassDev *cls = cls_g;
  nr_stage(sums, cls, base, P, nc, ns->n_paths);
  const int lane = threadIdx.x;
  auto member = [&](int k) { return cls[k].use && (mode == NRX_DERIV_TRUE || cls[k].has); };
  // derivatives at ns->x, mixed per partition as the host does
  double D1 = 0.0, D2 = 0.0;
  for (int p = 0; p < P; ++p) {
    auto d1 = [&](int k) {
      return cls[k].has && cls[k].path >= 0 ? sums[(static_cast<size_t>(cls[k].path) * 3 + 1) * P + p] : 0.0;
    };
    auto d2 = [&](int k) {
      return cls[k].has && cls[k].path >= 0 ? sums[(static_cast<size_t>(cls[k].path) * 3 + 2) * P + p] : 0.0;
    };
    int cnt = 0, last = -1;
    for (int k = lane; k < nc; k += 64)
      if (member(k)) {
        ++cnt;
        last = k;
      }
    const int n_members = nr_wave_isum(cnt), only = nr_wave_imax(last);
    double pd1 = 0.0, pd2 = 0.0;
    if (n_members == 1 && mode == NRX_DERIV_REFERENCE) {
      pd1 = d1(only);  // single-tree mode
      pd2 = d2(only);
    } else if (n_members > 0) {
      if (variant == NRX_AVERAGE_DISPLAYED_TREES) {
        double m = -INFINITY;
        for (int k = lane; k < nc; k += 64)
          if (member(k)) m = fmax(m, cls[k].logprob + nr_tree_logl(sums, cls[k], base, P, nc, k, p));
        const double mx = nr_wave_max(m);
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int k = lane; k < nc; k += 64) {
          if (!member(k)) continue;
          const double wgt = exp(cls[k].logprob + nr_tree_logl(sums, cls[k], base, P, nc, k, p) - mx);
          const double a1 = d1(k);
          s0 += wgt;
          s1 += wgt * a1;
          s2 += wgt * (a1 * a1 + d2(k));
        }
        s0 = nr_wave_sum(s0);
        s1 = nr_wave_sum(s1);
        s2 = nr_wave_sum(s2);
        pd1 = s1 / s0;
        pd2 = (s2 * s0 - s1 * s1) / (s0 * s0);
      } else {
        // the first class (lowest index) with the highest score, as the
        // host's strict comparison in class order picks it
        double best = -INFINITY;
        int kb = -1;
        for (int k = lane; k < nc; k += 64) {
          if (!member(k)) continue;
          const double ll = nr_tree_logl(sums, cls[k], base, P, nc, k, p);
          // reference mode: the tree maximising tree_logl * tree_prob
          // (LikelihoodDerivatives.cpp:161-167)
          const double score = mode == NRX_DERIV_REFERENCE ? ll * exp(cls[k].logprob) : cls[k].logprob + ll;
          if (score > best) {
            best = score;
            kb = k;
          }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
          const double ob = __shfl_xor(best, o);
          const int ok = __shfl_xor(kb, o);
          if (ok >= 0 && (kb < 0 || ob > best || (ob == best && ok < kb))) {
            best = ob;
            kb = ok;
          }
        }
        if (kb >= 0) {
          pd1 = d1(kb);
          pd2 = d2(kb);
        }
      }
    }
    D1 += pd1;
    D2 += pd2;
  }
  if (lane != 0) return;
  ns->d1 = D1;
  ns->d2 = D2;
  ns->iters += 1;
  if (ns->iters >= ns->max_iters) ns->done = 1;  // this step is the last
  // safeguarded Newton step: keep the maximiser's bracket [a, b], take the
  // Newton step when the curvature is negative and the step stays inside,
  // else grow geometrically toward a far bound or bisect
  const double x = ns->x, tol = ns->tol;
  double a = ns->a, b = ns->b;
  if (D1 > 0.0) {
    a = x;
  } else if (D1 < 0.0) {
    b = x;
  } else {
    ns->done = 1;
    return;
  }
  double nx = 0.0;
  bool ok = false;
  if (D2 < 0.0) {
    nx = x - D1 / D2;
    ok = a < nx && nx < b;
  }
  if (!ok) {
    if (D1 > 0.0 && b - x > 4.0 * (x + tol))
      nx = fmin(b, 2.0 * x + tol);
    else if (D1 < 0.0 && x - a > 4.0 * tol && x > 4.0 * a)
      nx = fmax(a, 0.5 * x);
    else
      nx = 0.5 * (a + b);
  }
  ns->a = a;
  ns->b = b;
  if (fabs(nx - x) < tol) ns->done = 1;
  ns->x = fmin(fmax(nx, ns->lo), ns->hi);
}

// network lnL at the final length vs at the start (both from the same tree
// lnLs); a worse length is given up (optimize_branch,
// BranchLengthOptimization.cpp:317-330)
__global__ void __launch_bounds__(64)
    k_newton_accept(const double *__restrict__ sums_g, int P, int nc,
                    const DerivClassDev *__restrict__ cls_g, const double *__restrict__ base_g,
                    int variant, NewtonState *__restrict__ ns) {
  const double *sums = sums_g, *base = base_g;
  const DerivClassDev *cls = cls_g;
  nr_stage(sums, cls, base, P, nc, ns->n_paths);
  const int lane = threadIdx.x;
  double l_new = 0.0, l_old = 0.0;
  for (int p = 0; p < P; ++p) {
    double mn = -INFINITY, mo = -INFINITY;
    for (int k = lane; k < nc; k += 64) {
      if (!cls[k].use) continue;
      mn = fmax(mn, cls[k].logprob + nr_tree_logl(sums, cls[k], base, P, nc, k, p));
      mo = fmax(mo, cls[k].logprob + base[static_cast<size_t>(p) * nc + k]);
    }
    mn = nr_wave_max(mn);
    mo = nr_wave_max(mo);
    if (variant == NRX_BEST_DISPLAYED_TREE) {
      l_new += mn;
      l_old += mo;
      continue;
    }
    double sn = 0.0, so = 0.0;
    for (int k = lane; k < nc; k += 64) {
      if (!cls[k].use) continue;
      sn += exp(cls[k].logprob + nr_tree_logl(sums, cls[k], base, P, nc, k, p) - mn);
      so += exp(cls[k].logprob + base[static_cast<size_t>(p) * nc + k] - mo);
    }
    l_new += mn + log(nr_wave_sum(sn));
    l_old += mo + log(nr_wave_sum(so));
  }
  if (lane != 0) return;
  ns->lnl_new = l_new;
  ns->lnl_old = l_old;
  ns->accepted = l_new >= l_old;
}

// The two kernels above in one, with no sumtable stored: each (path, site,
// category) walks its path adjoint and evaluates at b right away (the
// allocation-free variant of the branch derivatives re-walks the paths on
// every call instead of keeping sumtables).
template <int S, int C, typename T>
__global__ void __launch_bounds__(TILE *C)
    k_deriv_fused(const PathStep *__restrict__ steps, const int32_t *__restrict__ path_begin, int n_paths,
                  DevLayout L, const int32_t *__restrict__ tile_part,
                  const T *__restrict__ clv, const int32_t *__restrict__ scl, int scale_unit,
                  const typename TipType<S>::type *__restrict__ tips, const double *__restrict__ P,
                  const double *__restrict__ V, const double *__restrict__ Vinv,
                  const double *__restrict__ freqs, const double *__restrict__ ev,
                  const double *__restrict__ rates, const double *__restrict__ weights, double brlen,
                  double *__restrict__ partials) {
  constexpr int PW = TILE * C / 64;
  const int tile = blockIdx.x;
  const int part = tile_part[tile];
  const int c = threadIdx.x % C;
  const int64_t site = static_cast<int64_t>(tile) * TILE + threadIdx.x / C;
  const size_t n_wt = static_cast<size_t>(L.n_tiles) * PW;
  const int wt = tile * PW + threadIdx.x / 64;
  const double w = weights[site];
  double ex[S], lr[S];
  stage_exps<S, C>(ev, rates, part, brlen, c, ex, lr);
  double *sPp = dyn_lds<double>();
  for (int j = 0; j < n_paths; ++j) {
    double stv[S], v[3];
    const int b0 = path_begin[j], b1 = path_begin[j + 1];
    const int n_staged = S == 4 ? min(b1 - b0, PATH_STAGED_STEPS) : 0;
    if (S == 4 && j > 0) __syncthreads();  // every lane is done with the previous path's staged P
    stage_path_p<S, C>(sPp, steps + b0, n_staged, P, part, L);
    path_sumtable<S, C, T>(steps, b0, b1, sPp, n_staged, L, part, site, c, clv, scl, scale_unit, tips, P, V,
                           Vinv, freqs, stv);
    sumtable_partials<S, C>(stv, ex, lr, w, c, v);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
      for (int q = 0; q < 3; ++q) partials[(static_cast<size_t>(j) * 3 + q) * n_wt + wt] = v[q];
  }
}


// ============================================= batches of networks (FP32)
// blockIdx.y selects the network; every kernel reads that network's
// SmallDesc and does exactly what the single-network kernel does.
__global__ void k_pmatrices_multi(const SmallDesc *__restrict__ descs) {
  const SmallDesc &d = descs[blockIdx.y];
  const DevLayout L = d.L;
  const int S = L.S, C = L.C;
  const int64_t total = static_cast<int64_t>(L.n_parts) * L.n_edges * C * S * S;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int j = idx % S;
  const int i = (idx / S) % S;
  const int c = (idx / (S * S)) % C;
  const int e = (idx / (S * S * C)) % L.n_edges;
  const int p = idx / (static_cast<int64_t>(S) * S * C * L.n_edges);
  const double rt = d.rates[p * C + c];
  const double t = d.brlen[e];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int k = 0; k < S; ++k) {
    const double lr = d.ev[p * S + k] * rt;
    const double x = d.V[(p * S + i) * S + k] * exp(lr * t) * d.Vinv[(p * S + k) * S + j];
    a0 += x;
    a1 += x * lr;
    a2 += x * lr * lr;
  }
  d.P[idx] = a0;
  d.dP[idx] = a1;
  d.d2P[idx] = a2;
  d.Pf[idx] = static_cast<float>(a0);
}

__device__ void reduce_segments_body(const double *__restrict__ in, size_t row_stride, int pw,
                                     const int32_t *__restrict__ ptb, int P, int row, int p,
                                     double *__restrict__ out) {
  __shared__ double sh[4];
  const double *r = in + static_cast<size_t>(row) * row_stride;
  const int t0 = ptb[p] * pw, t1 = ptb[p + 1] * pw;
  double s = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += 256) s += r[t];
  s = block_sum<256>(s, sh);
  if (threadIdx.x == 0) out[static_cast<size_t>(row) * P + p] = s;
}

__global__ void __launch_bounds__(256) k_root_reduce_multi(const SmallDesc *__restrict__ descs) {
  const SmallDesc &d = descs[blockIdx.y];
  const int P = d.L.n_parts;
  if (static_cast<int>(blockIdx.x) >= d.n_root * P) return;
  reduce_segments_body(d.root_partials, static_cast<size_t>(d.L.n_tiles) * d.pw, d.pw,
                       d.part_tile_begin, P, blockIdx.x / P, blockIdx.x % P, d.root_sums);
}

// out[g][p] = sum of the gradient partials of row g over the site blocks of
// partition p and their waves (layout [block][row][wave], gbpt blocks per
// 64-site tile, gpw waves per block), in a fixed order
__device__ void reduce_grad_body(const double *__restrict__ in, int n_uses, int gpw, int gbpt,
                                 const int32_t *__restrict__ ptb, int P, int g, int p,
                                 double *__restrict__ out) {
  __shared__ double sh[4];
  const int64_t q1 = static_cast<int64_t>(ptb[p + 1] - ptb[p]) * gbpt * gpw;
  const int64_t b0 = static_cast<int64_t>(ptb[p]) * gbpt;
  double s = 0.0;
  for (int64_t q = threadIdx.x; q < q1; q += 256)
    s += in[grad_index(b0 + q / gpw, n_uses, g, gpw, static_cast<int>(q % gpw))];
  s = block_sum<256>(s, sh);
  if (threadIdx.x == 0) out[static_cast<size_t>(g) * P + p] = s;
}

__global__ void __launch_bounds__(256)
    k_reduce_grad(const double *__restrict__ in, int n_uses, int gpw, int gbpt,
                  const int32_t *__restrict__ ptb, int P, double *__restrict__ out) {
  reduce_grad_body(in, n_uses, gpw, gbpt, ptb, P, blockIdx.x / P, blockIdx.x % P, out);
}

__global__ void __launch_bounds__(256) k_use_reduce_multi(const SmallDesc *__restrict__ descs) {
  const SmallDesc &d = descs[blockIdx.y];
  const int P = d.L.n_parts;
  if (static_cast<int>(blockIdx.x) >= d.n_uses * P) return;
  reduce_grad_body(d.grad_partials, d.n_uses, d.gpw, d.gbpt, d.part_tile_begin, P, blockIdx.x / P,
                   blockIdx.x % P, d.use_sums);
}

__global__ void __launch_bounds__(256) k_mixture_multi(const SmallDesc *__restrict__ descs,
                                                       int variant) {
  const SmallDesc &d = descs[blockIdx.x];
  ResultLayout rl{};
  rl.total = d.rl_total;
  rl.per_part = d.rl_per_part;
  rl.root_logl = d.rl_root_logl;
  rl.tree_logl = d.rl_tree_logl;
  rl.alpha = d.rl_alpha;
  rl.dbrlen = d.rl_dbrlen;
  rl.part_dbrlen = d.rl_part_dbrlen;
  rl.dprob = d.rl_dprob;
  mixture_body(d.L, d.n_root, d.nc, d.R, variant, d.classes, d.class_dlogprob, d.root_sums,
               d.beta, d.res, rl);
}

__global__ void k_grad_total_multi(const SmallDesc *__restrict__ descs) {
  const SmallDesc &d = descs[blockIdx.y];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.L.n_edges) return;
  const int P = d.L.n_parts;
  double tot = 0.0;
  for (int p = 0; p < P; ++p) {
    double s = 0.0;
    for (int i = d.edge_use_begin[e]; i < d.edge_use_begin[e + 1]; ++i)
      s += d.use_sums[static_cast<size_t>(d.edge_use_list[i]) * P + p];
    d.res[d.rl_part_dbrlen + p * d.L.n_edges + e] = s;
    tot += s;
  }
  d.res[d.rl_dbrlen + e] = tot;
}

// Copies every network's result block next to each other for one D2H copy.
__global__ void k_gather_results(const SmallDesc *__restrict__ descs,
                                 const int64_t *__restrict__ offsets, double *__restrict__ out) {
  const SmallDesc &d = descs[blockIdx.y];
  const int64_t n = offsets[blockIdx.y + 1] - offsets[blockIdx.y];
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[offsets[blockIdx.y] + i] = d.res[i];
}

template <typename F>
void dispatch_sc(int S, int C, F &&f) {
  // instantiated shapes: DNA (4 states) and protein (20 states) with
  // 1, 2, 4 or 8 rate categories
  auto withC = [&](auto s_tag) {
    constexpr int SS = decltype(s_tag)::value;
    switch (C) {
      case 1: f(std::integral_constant<int, SS>{}, std::integral_constant<int, 1>{}); break;
      case 2: f(std::integral_constant<int, SS>{}, std::integral_constant<int, 2>{}); break;
      case 4: f(std::integral_constant<int, SS>{}, std::integral_constant<int, 4>{}); break;
      case 8: f(std::integral_constant<int, SS>{}, std::integral_constant<int, 8>{}); break;
      default: throw std::invalid_argument("unsupported number of rate categories");
    }
  };
  if (S == 4)
    withC(std::integral_constant<int, 4>{});
  else if (S == 20)
    withC(std::integral_constant<int, 20>{});
  else
    throw std::invalid_argument("unsupported number of states");
}

}  // namespace

void launch_pmatrices(Engine &e) {
  const int64_t total = static_cast<int64_t>(e.P) * e.E * e.C * e.S * e.S;
  const int nb = static_cast<int>((total + 255) / 256);
  hipLaunchKernelGGL(k_pmatrices, dim3(nb), dim3(256), 0, e.stream, e.L, e.d_V.as<double>(),
                     e.d_Vinv.as<double>(), e.d_eval.as<double>(), e.d_rates.as<double>(),
                     e.d_brlen.as<double>(), e.d_P.as<double>(), e.d_dP.as<double>(),
                     e.d_d2P.as<double>(), e.d_Pf.as<float>(), e.d_dPf.as<float>());
  NRX_HIP(hipGetLastError());
}

template <typename F>
void dispatch_c(int C, F &&f) {
  switch (C) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    default: throw std::invalid_argument("unsupported number of rate categories");
  }
}

void launch_forward_level(Engine &e, const DevOp *ops, int n_ops, const LevelChunk &) {
  if (n_ops == 0) return;
  if (e.S == AA_STATES) {
    dispatch_c(e.C, [&](auto c) {
      constexpr int C = decltype(c)::value;
      hipLaunchKernelGGL((k_forward_level_aa<C>), dim3(e.L.n_tiles_grid, n_ops), dim3(TILE * C), 0, e.stream,
                         ops, e.L, e.d_clv.as<double>(), e.d_scl.as<int32_t>(), e.d_tips.as<uint32_t>(),
                         e.d_P.as<double>(), e.d_tile_part.as<int32_t>());
    });
    NRX_HIP(hipGetLastError());
    return;
  }
  dispatch_sc(e.S, e.C, [&](auto s, auto c) {
    constexpr int S = decltype(s)::value, C = decltype(c)::value;
    constexpr int G = S == 4 ? PLAIN_G : 1;
    using TT = typename TipType<S>::type;
    hipLaunchKernelGGL((k_forward_level<S, C, G>), dim3(e.L.n_tiles_grid, (n_ops + G - 1) / G),
                       dim3(TILE * C), 0, e.stream, ops, n_ops, e.L, e.d_clv.as<double>(),
                       e.d_scl.as<int32_t>(), e.d_tips.as<TT>(), e.d_P.as<double>(),
                       e.d_tile_part.as<int32_t>());
  });
  NRX_HIP(hipGetLastError());
}

void launch_forward_walk(Engine &e, const DevOp *ops, const LevelChunk *walk, int n_walk) {
  if (e.S == AA_STATES) throw std::logic_error("launch_forward_walk: 20-state levels go through launch_forward_level");
  if (n_walk == 0) return;
  dispatch_sc(e.S, e.C, [&](auto s, auto c) {
    constexpr int S = decltype(s)::value, C = decltype(c)::value;
    constexpr int G = S == 4 ? PLAIN_G : 1;
    using TT = typename TipType<S>::type;
    if constexpr (S != AA_STATES)
      hipLaunchKernelGGL((k_forward_walk<S, C, G>), dim3(e.L.n_tiles_grid), dim3(TILE * C), 0, e.stream, ops, walk,
                         n_walk, e.L, e.d_clv.as<double>(), e.d_scl.as<int32_t>(), e.d_tips.as<TT>(),
                         e.d_P.as<double>(), e.d_tile_part.as<int32_t>());
  });
  NRX_HIP(hipGetLastError());
}

void launch_root_partials(Engine &e, bool seeds, const int32_t *rows, const int32_t *index, int n) {
  if (n == 0) return;
  dispatch_sc(e.S, e.C, [&](auto s, auto c) {
    constexpr int S = decltype(s)::value, C = decltype(c)::value;
    hipLaunchKernelGGL((k_root_partials<S, C>), dim3(e.L.n_tiles, n), dim3(TILE * C), 0, e.stream, rows, n, e.L,
                       e.d_clv.as<double>(), e.d_scl.as<int32_t>(), e.d_freqs.as<double>(),
                       e.d_weights.as<double>(), e.d_tile_part.as<int32_t>(), e.d_root_partials.as<double>(),
                       seeds ? e.d_site_seed.as<double>() : nullptr, index);
  });
  NRX_HIP(hipGetLastError());
}

void launch_root_partials(Engine &e, bool seeds) {
  launch_root_partials(e, seeds, e.d_root_slots.as<int32_t>(), nullptr, static_cast<int>(e.root_slots.size()));
}

void launch_mixture(Engine &e, int variant, int pw) {
  const int nr = static_cast<int>(e.root_slots.size());
  hipLaunchKernelGGL(k_reduce_segments, dim3(nr * e.P), dim3(256), 0, e.stream,
                     e.d_root_partials.as<double>(), static_cast<size_t>(e.L.n_tiles) * pw, pw,
                     e.d_part_tile_begin.as<int32_t>(), e.P, e.d_root_sums.as<double>());
  NRX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mixture, dim3(1), dim3(256), sizeof(double) * MIX_LDS_DOUBLES, e.stream, e.L, nr,
                     static_cast<int>(e.dag.classes.size()), e.R, variant,
                     e.d_classes.as<DevClass>(), e.d_class_dlogprob.as<double>(),
                     e.d_root_sums.as<double>(), e.d_beta.as<double>(), e.d_result.as<double>(),
                     e.rl);
  NRX_HIP(hipGetLastError());
}

void launch_adjoint_level(Engine &e, const LevelChunk &ch, const DevOp *ops) {
  const int n_ops = ch.op_end - ch.op_begin;
  if (n_ops == 0) return;
  if (e.S == AA_STATES) {
    dispatch_c(e.C, [&](auto c) {
      constexpr int C = decltype(c)::value;
      hipLaunchKernelGGL((k_adjoint_level_aa<C>), dim3(e.L.n_tiles_grid, n_ops), dim3(TILE * C), 0, e.stream,
                         ops + ch.op_begin, n_ops, e.L, e.d_clv.as<double>(), e.d_scl.as<int32_t>(),
                         e.d_adj.as<double>(), e.d_tips.as<uint32_t>(), e.d_P.as<double>(), e.d_Qd.as<double>(),
                         e.d_freqs.as<double>(), e.d_site_seed.as<double>(), e.d_tile_part.as<int32_t>(),
                         e.d_beta.as<double>(), static_cast<int>(e.root_slots.size()), e.n_uses,
                         e.d_grad_partials.as<double>());
    });
    NRX_HIP(hipGetLastError());
    return;
  }
  throw std::logic_error("launch_adjoint_level: 4-state sweeps go through launch_adjoint_levels");
}

void launch_adjoint_levels(Engine &e, const DevOp *ops, int c_hi, int c_lo) {
  const int nch = static_cast<int>(e.chunks.size());
  if (e.S == AA_STATES || c_lo < 0 || c_hi >= nch || c_lo > c_hi)
    throw std::logic_error("launch_adjoint_levels: 4-state level chunks expected");
  int widest = 0;
  for (int c = c_lo; c <= c_hi; ++c) widest = std::max(widest, e.chunks[c].op_end - e.chunks[c].op_begin);
  if (widest == 0) return;
  // the walk list holds the chunks top level first: chunk c at nch - 1 - c
  const LevelChunk *walk = e.d_adj_walk.as<LevelChunk>() + (nch - 1 - c_hi);
  const int n_walk = c_hi - c_lo + 1;
  dispatch_sc(e.S, e.C, [&](auto s, auto c) {
    constexpr int S = decltype(s)::value, C = decltype(c)::value;
    using TT = typename TipType<S>::type;
    const int max_staged = S == 4 ? std::min(widest, plain_staged_ops(C)) : 0;
    const size_t lds = S == 4 ? sizeof(double) * (max_staged * 2 + 1) * C * S * S : 0;
    if constexpr (S != AA_STATES)  // 20 states take k_adjoint_level_aa (refused above)
      hipLaunchKernelGGL((k_adjoint_level<S, C>), dim3(e.L.n_tiles_grid), dim3(TILE * C), lds, e.stream,
                         ops, walk, n_walk, max_staged, e.L, e.d_clv.as<double>(),
                         e.d_scl.as<int32_t>(), e.d_adj.as<double>(), e.d_tips.as<TT>(),
                         e.d_P.as<double>(), e.d_Qd.as<double>(), e.d_freqs.as<double>(),
                         e.d_site_seed.as<double>(), e.d_tile_part.as<int32_t>(), e.d_beta.as<double>(),
                         static_cast<int>(e.root_slots.size()), e.n_uses, e.d_grad_partials.as<double>());
  });
  NRX_HIP(hipGetLastError());
}

void launch_grad_reduce(Engine &e) {
  hipLaunchKernelGGL(k_reduce_grad, dim3(e.n_uses * e.P), dim3(256), 0, e.stream, e.d_grad_partials.as<double>(),
                     e.n_uses, e.gpw, e.gbpt, e.d_part_tile_begin.as<int32_t>(), e.P, e.d_use_sums.as<double>());
  NRX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_grad_total, dim3((e.E + 255) / 256), dim3(256), 0, e.stream, e.L,
                     e.d_edge_use_begin.as<int32_t>(), e.d_edge_use_list.as<int32_t>(),
                     e.d_use_sums.as<double>(), e.d_result.as<double>(), e.rl);
  NRX_HIP(hipGetLastError());
}



// Shapes of the 4-state persistent kernels: group width G in {2, 3, 4, 8}
// and TS in {16, 32, 64} sites per workgroup with TS C a multiple of 64.
template <typename F>
void dispatch_shape(int C, int G, int TS, F &&f) {
  auto with_ts = [&](auto c, auto g) {
    constexpr int CC = decltype(c)::value;
    if (TS == 64) return f(c, g, std::integral_constant<int, 64>{});
    if constexpr (32 * CC % 64 == 0)
      if (TS == 32) return f(c, g, std::integral_constant<int, 32>{});
    if constexpr (16 * CC % 64 == 0)
      if (TS == 16) return f(c, g, std::integral_constant<int, 16>{});
    throw std::invalid_argument("unsupported sites per workgroup for this number of categories");
  };
  dispatch_c(C, [&](auto c) {
    switch (G) {
      case 2: with_ts(c, std::integral_constant<int, 2>{}); break;
      case 3: with_ts(c, std::integral_constant<int, 3>{}); break;
      case 4: with_ts(c, std::integral_constant<int, 4>{}); break;
      case 8: with_ts(c, std::integral_constant<int, 8>{}); break;
      default: throw std::invalid_argument("unsupported group width");
    }
  });
}

void launch_forward_persistent(Engine &e, const NetDesc *nets) {
  launch_forward_persistent_multi(e.stream, e.C, e.tune_group, e.tune_tile,
                                  nets ? nets : e.d_netdesc.as<NetDesc>(), e.d_work.as<WorkItem>(),
                                  static_cast<int>(e.L.n_sites_pad / e.tune_tile),
                                  sizeof(float) * e.E * e.C * 16);
}

void launch_adjoint_persistent(Engine &e) {
  launch_adjoint_persistent_multi(e.stream, e.C, e.tune_group, e.tune_tile, e.tune_lanes,
                                  e.d_netdesc.as<NetDesc>(), e.d_work.as<WorkItem>(),
                                  static_cast<int>(e.L.n_sites_pad / e.tune_tile),
                                  adjoint_persistent_lds(e.E, e.C, e.tune_tile, e.pass_slots));
}

size_t adjoint_persistent_lds(int E, int C, int TS, int n_pass) {
  // P and Q tables, padded to 16 bytes, then n_pass slots of 16 bytes per
  // thread and 4 per site
  const size_t tables = (static_cast<size_t>(E) * C * 16 + C * 16 + 3) & ~size_t(3);
  return sizeof(float) * tables + static_cast<size_t>(n_pass) * TS * (C * 4 * sizeof(float) + sizeof(int));
}

// Pass slots for a reverse-sweep launch of n_work workgroups on the current
// device: the most slots (at most PASS_SLOTS) whose LDS still lets as many
// workgroups per CU be resident as the launch can use (its registers and
// tables allow, and its size fills); 0 when the tables alone do not fit.
int adjoint_pass_slots(int C, int G, int TS, int H, int E, int64_t n_work) {
  int dev = 0, n_cu = 0;
  NRX_HIP(hipGetDevice(&dev));
  NRX_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  int best = 0;
  dispatch_shape(C, G, TS, [&](auto c, auto g, auto ts) {
    constexpr int CC = decltype(c)::value, GG = decltype(g)::value, T = decltype(ts)::value;
    const void *kern = H == 2 ? reinterpret_cast<const void *>(&k_adjoint_split<CC, GG, T>)
                              : reinterpret_cast<const void *>(&k_adjoint_persistent<CC, GG, T>);
    auto resident = [&](int k) {
      int nb = 0;
      NRX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, T * CC * H,
                                                           adjoint_persistent_lds(E, C, TS, k)));
      return nb;
    };
    const int r0 = resident(0);
    if (r0 == 0) return;
    const int64_t want = std::min<int64_t>(r0, (n_work + n_cu - 1) / std::max(1, n_cu));
    for (int k = PASS_SLOTS; k > 0; --k)
      if (resident(k) >= want) {
        best = k;
        return;
      }
  });
  return best;
}

void launch_forward_persistent_multi(hipStream_t stream, int C, int G, int TS, const NetDesc *nets,
                                     const WorkItem *work, int n_work, size_t lds) {
  dispatch_shape(C, G, TS, [&](auto c, auto g, auto ts) {
    constexpr int CC = decltype(c)::value, GG = decltype(g)::value, T = decltype(ts)::value;
    hipLaunchKernelGGL((k_forward_persistent<CC, GG, T>), dim3(n_work), dim3(T * CC), lds, stream,
                       nets, work);
  });
  NRX_HIP(hipGetLastError());
}

void launch_adjoint_persistent_multi(hipStream_t stream, int C, int G, int TS, int H, const NetDesc *nets,
                                     const WorkItem *work, int n_work, size_t lds) {
  if (H != 1 && H != 2) throw std::invalid_argument("lanes per row must be 1 or 2");
  dispatch_shape(C, G, TS, [&](auto c, auto g, auto ts) {
    constexpr int CC = decltype(c)::value, GG = decltype(g)::value, T = decltype(ts)::value;
    if (H == 2)
      hipLaunchKernelGGL((k_adjoint_split<CC, GG, T>), dim3(n_work), dim3(T * CC * 2), lds, stream, nets,
                         work);
    else
      hipLaunchKernelGGL((k_adjoint_persistent<CC, GG, T>), dim3(n_work), dim3(T * CC), lds, stream,
                         nets, work);
  });
  NRX_HIP(hipGetLastError());
}

void launch_mfma_tables(Engine &e) {
  const int64_t total = static_cast<int64_t>(e.P) * e.E * e.C * AA_KINDS * AA_TABLE;
  const int nb = static_cast<int>((total + 255) / 256);
  hipLaunchKernelGGL(k_mfma_tables, dim3(nb), dim3(256), 0, e.stream, e.L, e.d_P.as<double>(),
                     e.d_Pa.as<float>());
  NRX_HIP(hipGetLastError());
}

// one workgroup per 32-site tile: the work table holds 2 n_tiles items
void launch_forward_mfma(Engine &e, const NetDesc *nets) {
  dispatch_c(e.C, [&](auto c) {
    constexpr int CC = decltype(c)::value;
    auto go = [&](auto g) {
      constexpr int GG = decltype(g)::value;
      hipLaunchKernelGGL((k_forward_mfma<CC, GG>), dim3(2 * e.L.n_tiles), dim3(64 * CC), 0,
                         e.stream, nets ? nets : e.d_netdesc.as<NetDesc>(), e.d_work.as<WorkItem>());
    };
    switch (e.tune_group) {
      case 1: go(std::integral_constant<int, 1>{}); break;
      case 2: go(std::integral_constant<int, 2>{}); break;
      case 4: go(std::integral_constant<int, 4>{}); break;
      default: throw std::invalid_argument("unsupported group width for 20-state data");
    }
  });
  NRX_HIP(hipGetLastError());
}

void launch_adjoint_mfma(Engine &e) {
  dispatch_c(e.C, [&](auto c) {
    constexpr int CC = decltype(c)::value;
    hipLaunchKernelGGL((k_adjoint_mfma<CC>), dim3(2 * e.L.n_tiles), dim3(64 * CC), 0, e.stream,
                       e.d_netdesc.as<NetDesc>(), e.d_work.as<WorkItem>());
  });
  NRX_HIP(hipGetLastError());
}

void launch_sumtables(Engine &e, int n_paths, int max_steps) {
  if (n_paths == 0) return;
  const size_t lds = path_lds_bytes(e.S, e.C, max_steps);
  dispatch_sc(e.S, e.C, [&](auto s, auto c) {
    constexpr int S = decltype(s)::value, C = decltype(c)::value;
    using TT = typename TipType<S>::type;
    if (e.opt.precision == NRX_FP32) {
      hipLaunchKernelGGL((k_sumtables<S, C, float>), dim3(e.L.n_tiles, n_paths), dim3(TILE * C), lds,
                         e.stream, e.d_steps.as<PathStep>(), e.d_path_begin.as<int32_t>(), n_paths,
                         e.L, e.d_tile_part.as<int32_t>(),
                         e.d_clv.as<float>(), e.d_scl.as<int32_t>(), 1, e.d_tips.as<TT>(),
                         e.d_P.as<double>(), e.d_V.as<double>(), e.d_Vinv.as<double>(),
                         e.d_freqs.as<double>(), e.d_st.as<double>());
    } else {
      hipLaunchKernelGGL((k_sumtables<S, C, double>), dim3(e.L.n_tiles, n_paths), dim3(TILE * C), lds,
                         e.stream, e.d_steps.as<PathStep>(), e.d_path_begin.as<int32_t>(), n_paths,
                         e.L, e.d_tile_part.as<int32_t>(),
                         e.d_clv.as<double>(), e.d_scl.as<int32_t>(), 256, e.d_tips.as<TT>(),
                         e.d_P.as<double>(), e.d_V.as<double>(), e.d_Vinv.as<double>(),
                         e.d_freqs.as<double>(), e.d_st.as<double>());
    }
  });
  NRX_HIP(hipGetLastError());
}

void launch_sumtable_eval(Engine &e, int n_paths, double brlen) {
  if (n_paths == 0) return;
  dispatch_sc(e.S, e.C, [&](auto s, auto c) {
    constexpr int S = decltype(s)::value, C = decltype(c)::value;
    hipLaunchKernelGGL((k_sumtable_eval<S, C>), dim3(e.L.n_tiles, n_paths), dim3(TILE * C), 0, e.stream,
                       e.d_st.as<double>(), n_paths, e.L, e.d_tile_part.as<int32_t>(),
                       e.d_eval.as<double>(), e.d_rates.as<double>(), e.d_weights.as<double>(),
                       brlen, e.d_st_partials.as<double>());
  });
  NRX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_segments, dim3(n_paths * 3 * e.P), dim3(256), 0, e.stream,
                     e.d_st_partials.as<double>(), static_cast<size_t>(e.L.n_tiles) * e.pw, e.pw,
                     e.d_part_tile_begin.as<int32_t>(), e.P, e.d_st_sums.as<double>());
  NRX_HIP(hipGetLastError());
}


// Scratch of the allocation-free branch derivatives (nrx_brlen_derivatives_noalloc):
// a fixed device array, one copy per GPU, whose address host code takes
// once per device from hipGetSymbolAddress (the array's name in host code
// is not a device address).
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

void *noalloc_scratch() {
  static std::mutex mu;
  static std::map<int, void *> ptr_of_device;
  int dev = 0;
  NRX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  const auto it = ptr_of_device.find(dev);
  if (it != ptr_of_device.end()) return it->second;
  void *ptr = nullptr;
  NRX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) throw std::runtime_error("HIP error: no device address for the no-alloc scratch");
  ptr_of_device[dev] = ptr;
  return ptr;
}

void launch_deriv_fused(Engine &e, const PathStep *steps, const int32_t *path_begin, int n_paths, int max_steps,
                        double brlen, double *partials, double *sums) {
  if (n_paths == 0) return;
  const size_t lds = path_lds_bytes(e.S, e.C, max_steps);
  dispatch_sc(e.S, e.C, [&](auto s, auto c) {
    constexpr int S = decltype(s)::value, C = decltype(c)::value;
    using TT = typename TipType<S>::type;
    if (e.opt.precision == NRX_FP32)
      hipLaunchKernelGGL((k_deriv_fused<S, C, float>), dim3(e.L.n_tiles), dim3(TILE * C), lds, e.stream, steps,
                         path_begin, n_paths, e.L, e.d_tile_part.as<int32_t>(), e.d_clv.as<float>(),
                         e.d_scl.as<int32_t>(), 1, e.d_tips.as<TT>(), e.d_P.as<double>(), e.d_V.as<double>(),
                         e.d_Vinv.as<double>(), e.d_freqs.as<double>(), e.d_eval.as<double>(),
                         e.d_rates.as<double>(), e.d_weights.as<double>(), brlen, partials);
    else
      hipLaunchKernelGGL((k_deriv_fused<S, C, double>), dim3(e.L.n_tiles), dim3(TILE * C), lds, e.stream, steps,
                         path_begin, n_paths, e.L, e.d_tile_part.as<int32_t>(),
                         e.d_clv.as<double>(), e.d_scl.as<int32_t>(), 256, e.d_tips.as<TT>(),
                         e.d_P.as<double>(), e.d_V.as<double>(), e.d_Vinv.as<double>(), e.d_freqs.as<double>(),
                         e.d_eval.as<double>(), e.d_rates.as<double>(), e.d_weights.as<double>(), brlen,
                         partials);
  });
  NRX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_segments, dim3(n_paths * 3 * e.P), dim3(256), 0, e.stream, partials,
                     static_cast<size_t>(e.L.n_tiles) * e.pw, e.pw, e.d_part_tile_begin.as<int32_t>(), e.P, sums);
  NRX_HIP(hipGetLastError());
}

void launch_model_grad(Engine &e, int n_split, double *g_out, double *r_out) {
  if (e.S != 4) throw std::invalid_argument("model gradients are built for 4-state data");
  dispatch_c(e.C, [&](auto c) {
    constexpr int C = decltype(c)::value;
    hipLaunchKernelGGL((k_model_grad<4, C>), dim3(static_cast<unsigned>(e.ops_all.size()), e.P * n_split),
                       dim3(TILE * C), 0, e.stream, e.d_ops.as<DevOp>(), e.L, e.d_clv.as<double>(),
                       e.d_scl.as<int32_t>(), e.d_adj.as<double>(), e.d_tips.as<uint8_t>(), e.d_P.as<double>(),
                       e.d_freqs.as<double>(), e.d_weights.as<double>(), e.d_beta.as<double>(),
                       static_cast<int>(e.root_slots.size()), e.d_part_tile_begin.as<int32_t>(), n_split, e.n_uses,
                       g_out, r_out);
  });
  NRX_HIP(hipGetLastError());
}

void enqueue_newton(Engine &e, int variant, int mode, int rounds, int max_paths, NewtonState *st,
                    const DerivClassDev *cls, const double *base_logl) {
  const int nc = static_cast<int>(e.dag.classes.size());
  const size_t row_stride = static_cast<size_t>(e.L.n_tiles) * e.pw;
  auto eval_reduce = [&](int final_eval) {
    dispatch_sc(e.S, e.C, [&](auto s, auto c) {
      constexpr int S = decltype(s)::value, C = decltype(c)::value;
      hipLaunchKernelGGL((k_sumtable_eval_nr<S, C>), dim3(e.L.n_tiles, max_paths), dim3(TILE * C), 0, e.stream,
                         e.d_st.as<double>(), e.L, e.d_tile_part.as<int32_t>(), e.d_eval.as<double>(),
                         e.d_rates.as<double>(), e.d_weights.as<double>(), st, final_eval,
                         e.d_st_partials.as<double>());
    });
    NRX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_segments_nr, dim3(max_paths * 3 * e.P), dim3(256), 0, e.stream,
                       e.d_st_partials.as<double>(), row_stride, e.pw, e.d_part_tile_begin.as<int32_t>(), e.P, st,
                       final_eval, e.d_st_sums.as<double>());
    NRX_HIP(hipGetLastError());
  };
  for (int it = 0; it < rounds; ++it) {
    eval_reduce(0);
    hipLaunchKernelGGL(k_newton_step, dim3(1), dim3(64), NR_LDS_BYTES, e.stream, e.d_st_sums.as<double>(), e.P,
                       nc, cls, base_logl, variant, mode, st);
    NRX_HIP(hipGetLastError());
  }
  eval_reduce(1);
  hipLaunchKernelGGL(k_newton_accept, dim3(1), dim3(64), NR_LDS_BYTES, e.stream, e.d_st_sums.as<double>(), e.P,
                     nc, cls, base_logl, variant, st);
  NRX_HIP(hipGetLastError());
}

void launch_pmatrices_multi(hipStream_t s, const SmallDesc *d, int n, int max_elems) {
  hipLaunchKernelGGL(k_pmatrices_multi, dim3((max_elems + 255) / 256, n), dim3(256), 0, s, d);
  NRX_HIP(hipGetLastError());
}

void launch_root_reduce_multi(hipStream_t s, const SmallDesc *d, int n, int max_rows) {
  hipLaunchKernelGGL(k_root_reduce_multi, dim3(max_rows, n), dim3(256), 0, s, d);
  NRX_HIP(hipGetLastError());
}

void launch_mixture_multi(hipStream_t s, const SmallDesc *d, int n, int variant) {
  hipLaunchKernelGGL(k_mixture_multi, dim3(n), dim3(256), sizeof(double) * MIX_LDS_DOUBLES, s, d, variant);
  NRX_HIP(hipGetLastError());
}

void launch_grad_reduce_multi(hipStream_t s, const SmallDesc *d, int n, int max_rows,
                              int max_edges) {
  hipLaunchKernelGGL(k_use_reduce_multi, dim3(max_rows, n), dim3(256), 0, s, d);
  NRX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_grad_total_multi, dim3((max_edges + 255) / 256, n), dim3(256), 0, s, d);
  NRX_HIP(hipGetLastError());
}

void launch_gather_results(hipStream_t s, const SmallDesc *d, int n, const int64_t *offsets,
                           double *out, int max_size) {
  hipLaunchKernelGGL(k_gather_results, dim3((max_size + 255) / 256, n), dim3(256), 0, s, d,
                     offsets, out);
  NRX_HIP(hipGetLastError());
}

}  // namespace nrx
