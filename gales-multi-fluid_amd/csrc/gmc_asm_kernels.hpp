// gmc_asm_kernels.hpp -- element assembly kernels for gfx950 (instantiated per
// element shape in gmc_asm_<dim>d<nc>.hip, dispatched from gmc_assemble.hip).
//
// One element per 4-lane group ("quad"); lane q of the quad evaluates
// interior quadrature point q (fluid_2d.hpp:177-199 loops the points of one
// element serially).  Triangles have 3 points, so lane 3 duplicates point 2
// with weight 0; tetrahedra use all 4 lanes.  Properties, flux matrices and
// tau are evaluated in FP64; the element residual is FP64; the element matrix
// is formed in TM: double for the plain form, float for the FP32 tier.  Each
// lane produces one ndof x ndof block of its point's contribution at a time,
// the quad sums it with DPP quad permutes, and the four lanes store disjoint
// entries.
//
// Scatter (scatter.hpp:37-98) is deterministic without atomics, in one of two
// ways:
//  - element buffer (default): every element writes its Dirichlet-processed
//    blocks (TM) and rhs part (FP64) to a private slot (stores only, one launch
//    over all elements); boundary integrals add into their own element's slot;
//    gather kernels then sum each BSR block's and each node's contributions in
//    fixed element order, in FP64.
//  - matrix-free products (MODE_JVP, MODE_VJPT): every element forms its FP64
//    matrix on the fly and stores m_e x_e (Dirichlet rows: avg_val x), or the
//    transpose product, in the rhs part of its slot; the node gather then
//    yields y = J x or y = J^T x without storing J.
//  - colour scatter (when the buffer does not fit): elements are launched one
//    colour at a time and read-modify-write the FP64 BSR values; no two
//    elements of a colour share a node, so each entry has one writer per
//    launch.
// The Dirichlet rule (avg_val = (1 + sum_i m_ii)/n on blocked diagonals, zero
// rows, zero rhs) is applied per element and per integral (interior,
// boundary), as the reference applies it per scatter call.
#pragma once

#include <hip/hip_runtime.h>

#include "gmc_element.hpp"
#include "gmc_kernels.hpp"
#include "gmc_wave.hpp"

namespace gmc {

// Minimum waves per SIMD the element kernels are compiled for (measured on
// MI355X).  The FP64 element-matrix phase keeps ~450 values per lane live
// (point record, tau D, T_a, D T_a, one block): at one wave per SIMD it fits
// the unified register file (VGPRs + AGPRs, 16 B of scratch); forced to two
// waves it spills 756 B per lane and runs 1.6x slower.  The residual-only and
// FP32-matrix 2D kernels fit two waves with little scratch and gain from the
// latency hiding (2D channel of 4.19M triangles: residual 2.38 -> 1.62 ms,
// FP32-tier R+J 7.20 -> 4.74 ms).  Three waves spill heavily everywhere.  The
// 3D kernels (ndof 7, 4 nodes) spill at one wave already (1 KB per lane in
// FP64 with the rolled column-node loop); at two the FP32-tier kernel on the
// 128^3 Kuhn box takes 96 ms instead of 47, so they stay at one.

// v[i] for a lane-dependent i < NV, by selects: a runtime index into a
// private array would place the array in scratch memory
template <int NV, typename T>
__device__ __forceinline__ T pick(const T* v, int i) {
  T x = v[0];
  GMC_UNROLL for (int k = 1; k < NV; k++) x = (i == k) ? v[k] : x;
  return x;
}

// this point's record in a magma point table (states records of NC
// components each), or null without magma components
template <int NC>
__device__ __forceinline__ const double* magma_row(const double* mt, long long point, int states) {
  return mt ? mt + point * states * NC * kMagmaVals : nullptr;
}

template <int DIM, bool MAT, typename TM>
constexpr int waves_per_simd() {
  return (DIM == 3 || (MAT && sizeof(TM) == 8)) ? 1 : 2;
}

// ---------------------------------------------------------------------------
// interior elements (all of them, or one colour in MODE_SCATTER)
// ---------------------------------------------------------------------------
template <int DIM, int NC, bool MAT, bool RHS, int MODE, typename TM>
__global__ __launch_bounds__(256, (waves_per_simd<DIM, MAT, TM>())) void k_interior(const AsmArgs a) {
  using S = Shape<DIM, NC>;
  using L = EbLayout<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE, NB = N * N;
  constexpr bool OUT = MODE == MODE_DENSE;
  constexpr bool TR = MODE == MODE_VJPT;
  constexpr bool JVP = MODE == MODE_JVP || TR;
  constexpr bool EBUF = MODE == MODE_EBUF || JVP;
  const int q = threadIdx.x & 3;
  const long long gid = (long long)blockIdx.x * (blockDim.x >> 2) + (threadIdx.x >> 2);
  if (gid >= a.n_items) return;  // whole quads leave together
  const long long e = a.items[gid];
  // element-buffer slots of the element's blocks, loaded before the physics
  int eslot[NEN * NEN];
  GMC_UNROLL for (int k = 0; k < NEN * NEN; k++)
    eslot[k] = (MODE == MODE_EBUF && MAT) ? a.slot[e * NEN * NEN + k] : 0;
  const DevParams& P = *a.params;
  Material<double> M;
  load_material<double, NC>(P, M);

  int en[NEN];
  double xe[NEN * DIM], Ie[NE], Pe[NE];
  gather<DIM, NC>(a, e, en, xe, Ie, Pe);
  double xv[JVP ? NE : 1], yv[JVP ? NE : 1];
  if constexpr (JVP) {
    GMC_UNROLL for (int n = 0; n < NEN; n++)
      GMC_UNROLL for (int k = 0; k < N; k++) {
        xv[n * N + k] = a.xvec[(long long)en[n] * N + k];
        yv[n * N + k] = 0.0;
      }
  }
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) {
    if (q == 0) atomicOr(a.status, 1);
    return;
  }
  Point<double, DIM, NC> pt;
  const int qp = q < S::NQ ? q : S::NQ - 1;
  const int st = point_setup<double, DIM, NC>(P, M, g, qp, q < S::NQ ? 1.0 : 0.0, Ie, Pe, pt,
                                              magma_row<NC>(a.mt, e * S::NQ + qp, 2));
  if (st && q < S::NQ) atomicOr(a.status, st);

  unsigned dm[NEN];
  GMC_UNROLL for (int n = 0; n < NEN; n++) dm[n] = OUT ? 0u : a.dmask[en[n]];

  if (RHS) {
    double r[NE];
    point_residual<double, DIM, NC>(P, pt, g, r);
    double* ebr = EBUF ? a.ebr + e * L::RS : nullptr;
    double rs[(NE + 3) / 4];
    quad_rs<double, NE>(r, q, rs);  // rs[k]: entry 4k + q
    GMC_UNROLL for (int k = 0; k < (NE + 3) / 4; k++) {
      const int t = 4 * k + q;
      if (t >= NE) continue;
      const int b = t / N, jb = t - b * N;
      const double rr = rs[k];
      const bool blocked = (pick<NEN>(dm, b) >> jb) & 1u;
      if (OUT) a.out_r[gid * NE + t] = -rr;
      else if (EBUF) ebr[t] = blocked ? 0.0 : -rr;
      else if (!blocked) a.rhs[(long long)pick<NEN>(en, b) * N + jb] -= rr;
    }
  }

  if (MAT) {
    MatPoint<TM, DIM, NC> mp;
    make_matpoint<TM, DIM, NC>(P, pt, g, mp);
    TM* ebm = EBUF ? reinterpret_cast<TM*>(a.ebm) : nullptr;
    // matrix-free modes sum every entry into every lane (quad_sum); the
    // storing modes reduce-scatter, each lane keeping entries 4k + q, and the
    // trace is this lane's diagonal entries, quad-summed after the blocks
    TM trace = TM(1.0), tpart = TM(0.0);
    // rolled where measured faster: the 3D FP64 matrix-free modes, and the
    // storing modes with 7 dofs in both precisions (FP32 tier: 47.2 vs
    // 48.3 ms on 12.6M tetrahedra); with 8 dofs the FP64 storing modes run
    // faster unrolled (2.65M tetrahedra: element buffer 41.0 vs 43.9 ms,
    // colour scatter 61.3 vs 67.6 ms)
    constexpr bool ROLLED = DIM == 3 && (sizeof(TM) == 8 ? (JVP || N == 7) : (!JVP && N == 7));
    if constexpr (ROLLED) {
      // 3D: the column-node loop is a loop, not unrolled, so that only one
      // column node's T_a, D T_a and blocks are live.  The shape data (and in
      // the matrix-free modes the node blocks of x and y) rotate by one node
      // per iteration, so column node ia is always node 0 and node bb is node
      // (bb + ia) mod 4: every index into a private array stays a compile-time
      // constant.  After the 4 iterations the rotation is back where it began.
      static_assert(NEN == 4, "rotation by node index mod 4");
      MatPhase<TM, DIM, NC> ph2;
      ph2.init(mp);
#pragma clang loop unroll(disable)
      for (int ia = 0; ia < NEN; ia++) {
        ph2.col(mp, 0);
        GMC_UNROLL for (int bb = 0; bb < NEN; bb++) {
          const int ib = (bb + ia) & 3;
          TM m[N][N];
          ph2.block(mp, bb, 0, m);
          const unsigned dmb = pick<NEN>(dm, ib);
          if constexpr (JVP) {
            GMC_UNROLL for (int jb = 0; jb < N; jb++)
              GMC_UNROLL for (int ja = 0; ja < N; ja++) {
                const TM mm = quad_sum(m[jb][ja]);
                if (bb == 0 && jb == ja) trace += mm;
                if constexpr (TR) {
                  // blocked rows of m are replaced by the avg_val diagonal below
                  if (!((dmb >> jb) & 1u)) yv[ja] += (double)mm * xv[bb * N + jb];
                } else {
                  yv[bb * N + jb] += (double)mm * xv[ja];
                }
              }
          } else {
            const int sl = (MODE == MODE_EBUF) ? pick<NEN * NEN>(eslot, ib * NEN + ia) : 0;
            const long long blk = (OUT || EBUF) ? 0 : a.blkmap[e * NEN * NEN + ib * NEN + ia];
            TM mv[(NB + 3) / 4];  // this lane's entries t = 4k + q of the block
            quad_rs<TM, NB>(&m[0][0], q, mv);
            GMC_UNROLL for (int k = 0; k < (NB + 3) / 4; k++) {
              const int t = 4 * k + q, jb = t / N, ja = t - jb * N;
              if (t >= NB) continue;
              if (bb == 0 && jb == ja) tpart += mv[k];
              const TM mm = (!OUT && ((dmb >> jb) & 1u)) ? TM(0) : mv[k];
              if (OUT)
                a.out_m[gid * NE * NE + (long long)(ib * N + jb) * NE + ia * N + ja] = (double)mm;
              else if (EBUF) ebm[(long long)sl * NB + t] = mm;
              else a.vals[blk * NB + t] += (double)mm;
            }
          }
        }
        // node ia + 1 becomes node 0
        const TM s0 = mp.sh[0];
        GMC_UNROLL for (int n = 0; n < NEN - 1; n++) mp.sh[n] = mp.sh[n + 1];
        mp.sh[NEN - 1] = s0;
        GMC_UNROLL for (int k = 0; k < DIM; k++) {
          const TM d0 = mp.dsh[k][0];
          GMC_UNROLL for (int n = 0; n < NEN - 1; n++) mp.dsh[k][n] = mp.dsh[k][n + 1];
          mp.dsh[k][NEN - 1] = d0;
        }
        if constexpr (JVP) {
          GMC_UNROLL for (int k = 0; k < N; k++) {
            const double x0 = xv[k], y0 = yv[k];
            GMC_UNROLL for (int n = 0; n < NEN - 1; n++) {
              xv[n * N + k] = xv[(n + 1) * N + k];
              yv[n * N + k] = yv[(n + 1) * N + k];
            }
            xv[(NEN - 1) * N + k] = x0;
            yv[(NEN - 1) * N + k] = y0;
          }
        }
      }
    } else {
    MatPhase<TM, DIM, NC> ph2;
    ph2.init(mp);
    GMC_UNROLL for (int ia = 0; ia < NEN; ia++) {
      ph2.col(mp, ia);
      GMC_UNROLL for (int ib = 0; ib < NEN; ib++) {
        TM m[N][N];
        ph2.block(mp, ib, ia, m);
        if constexpr (JVP) {
          GMC_UNROLL for (int jb = 0; jb < N; jb++)
            GMC_UNROLL for (int ja = 0; ja < N; ja++) {
              const TM mm = quad_sum(m[jb][ja]);
              if (ib == ia && jb == ja) trace += mm;
              if constexpr (TR) {
                // blocked rows of m are replaced by the avg_val diagonal below
                if (!((dm[ib] >> jb) & 1u)) yv[ia * N + ja] += (double)mm * xv[ib * N + jb];
              } else {
                yv[ib * N + jb] += (double)mm * xv[ia * N + ja];
              }
            }
        } else {
          const long long blk = (OUT || EBUF) ? 0 : a.blkmap[e * NEN * NEN + ib * NEN + ia];
          TM mv[(NB + 3) / 4];  // this lane's entries t = 4k + q of the block
          quad_rs<TM, NB>(&m[0][0], q, mv);
          GMC_UNROLL for (int k = 0; k < (NB + 3) / 4; k++) {
            const int t = 4 * k + q, jb = t / N, ja = t - jb * N;
            if (t >= NB) continue;
            if (ib == ia && jb == ja) tpart += mv[k];
            const TM mm = (!OUT && ((dm[ib] >> jb) & 1u)) ? TM(0) : mv[k];
            if (OUT)
              a.out_m[gid * NE * NE + (long long)(ib * N + jb) * NE + ia * N + ja] = (double)mm;
            else if (EBUF) ebm[(long long)eslot[ib * NEN + ia] * NB + t] = mm;
            else a.vals[blk * NB + t] += (double)mm;
          }
        }
      }
    }
    }
    if constexpr (!JVP) trace += quad_sum(tpart);
    if constexpr (JVP) {
      // rows blocked by the Dirichlet rule carry avg_val on the diagonal only
      const double avg = (double)(trace / TM(NE));
      double* ebr = a.ebr + e * L::RS;
      GMC_UNROLL for (int t = 0; t < NE; t++)
        if ((t & 3) == q) {
          const bool blocked = (dm[t / N] >> (t % N)) & 1u;
          if (TR) ebr[t] = yv[t] + (blocked ? avg * xv[t] : 0.0);
          else ebr[t] = blocked ? avg * xv[t] : yv[t];
        }
    } else if (!OUT) {
      // Dirichlet rows: diagonal = avg_val (scatter.hpp:65-78).  In the
      // element buffer the same lane overwrites the 0 it stored above.
      const TM avg = trace / TM(NE);
      GMC_UNROLL for (int ib = 0; ib < NEN; ib++) {
        if (!dm[ib]) continue;
        GMC_UNROLL for (int jb = 0; jb < N; jb++) {
          const int t = jb * N + jb;
          if ((t & 3) == q && ((dm[ib] >> jb) & 1u)) {
            if (EBUF) ebm[(long long)eslot[ib * NEN + ib] * NB + t] = avg;
            else a.vals[(long long)a.blkmap[e * NEN * NEN + ib * NEN + ib] * NB + t] += (double)avg;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// FP32-tier element matrices with the least-squares term on the matrix cores
// (element-buffer mode; triangles with 2 components: 3 nodes x 5 dofs, 15
// rows and columns).  The least-squares block (b, a) of an element is
//   sum_q sum_l L_b(q)[i][l] T_a(q)[l][j],   T_a = JxW tau R_a
// (first_part_b T_a, fluid_2d.hpp:279-305; L_b's columns: lsq_left_col), one
// product L R per element with L[(b,i)][(l,q)] and R[(l,q)][(a,j)].  The
// inner index runs l-major over the quad's 4 lanes, so that inner slice l is
// one v_mfma_f32_16x16x4f32 whose 4 inner entries are the quad's 4 points
// (lane 3 repeats point 2 with weight 0: its T is 0, its L finite).  Per
// slice the quads write their operand entries to LDS for all 16 elements of
// the wave, then every lane reads its A and B entries of each element
// (lane k*16 + m supplies A[m][k] and B[k][m]; its result entries are
// D[4 (lane/16) + r][lane % 16], r = 0..3).  The 16 products come back
// through LDS to the quad layout, where the Galerkin, diffusion and DC terms
// (point_block2_full without least squares, VALU) are quad-reduced and
// added before the stores of the element-buffer mode.  Every lane runs the
// whole kernel (the MFMAs and barriers need them all); lanes past the last
// element, or of an element whose geometry failed, store nothing.  The
// staging areas are per wave, so the waves synchronise only within
// themselves.
// ---------------------------------------------------------------------------
namespace mfma {
constexpr int KS = 260;        // floats per inner index k of a staged operand (16 x 16 + 4)
constexpr int OPS = 4 * KS;    // one operand of one wave
constexpr int WAVE = 2 * OPS;  // A and B of one wave (8,320 B); also holds 8 elements' 16 x 16 results
}  // namespace mfma
static_assert(mfma::WAVE * sizeof(float) == PATCH_MFMA_WAVE_BYTES, "patch staging size");

typedef float gmc_f4 __attribute__((ext_vector_type(4)));

// LDS writes of this wave complete and visible to all its lanes (the staging
// areas are private to a wave: no workgroup barrier)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The least-squares products of the wave's 16 elements on the matrix cores,
// returned in the quad layout: lsq[b][a][k] = entry 4k + q of block (b, a) of
// this lane's element.  sw: the wave's staging area (mfma::WAVE floats of
// LDS); every lane of the wave takes part.
template <int DIM, int NC>
__device__ __forceinline__ void lsq_mfma(const MatPoint<float, DIM, NC>& mp, float* sw, int lane,
                                         float (*lsq)[DIM + 1][(Shape<DIM, NC>::NDOF * Shape<DIM, NC>::NDOF + 3) / 4]) {
  using S = Shape<DIM, NC>;
  using TM = float;
  constexpr int N = S::NDOF, NEN = S::NEN, NB = N * N;
  static_assert(DIM == 2 && NEN * N <= 16, "one 16 x 16 tile per element");
  const int q = lane & 3, ew = lane >> 2;
  // least squares on the matrix cores: acc[w] = L R of the wave's element w
  gmc_f4 acc[16];
  GMC_UNROLL for (int w = 0; w < 16; w++) acc[w] = gmc_f4{0.f, 0.f, 0.f, 0.f};
  {
    MatPrep<TM, DIM, NC> pr;
    point_prep_full<TM, DIM, NC>(mp, pr);
    float* sa = sw;               // A^T: [k][m][w]
    float* sb = sw + mfma::OPS;   // B:   [k][n][w]
    const int kk = lane >> 4, mm = lane & 15;
    GMC_UNROLL for (int l = 0; l < N; l++) {
      // this point's inner entry (l, q): row l of T_a for every a, column l
      // of L_b for every b
      TM col[NEN][N], row[NEN][N];
      GMC_UNROLL for (int b = 0; b < NEN; b++) lsq_left_col<TM, DIM, NC>(mp, b, l, col[b]);
      GMC_UNROLL for (int c = 0; c < NEN; c++) {
        TM T[N][N], s_[N], eT[N];
        point_col<TM, DIM, NC>(mp, pr, c, T, s_, eT);
        GMC_UNROLL for (int j = 0; j < N; j++) row[c][j] = T[l][j];
      }
      wave_lds_sync();  // the previous slice's reads are done
      GMC_UNROLL for (int m = 0; m < 16; m++) {
        const int b = m / N, i = m - b * N;
        sa[q * mfma::KS + m * 16 + ew] = m < NEN * N ? col[b < NEN ? b : 0][i] : 0.f;
        sb[q * mfma::KS + m * 16 + ew] = m < NEN * N ? row[b < NEN ? b : 0][i] : 0.f;
      }
      wave_lds_sync();
      const float4* pa = reinterpret_cast<const float4*>(sa + kk * mfma::KS + mm * 16);
      const float4* pb = reinterpret_cast<const float4*>(sb + kk * mfma::KS + mm * 16);
      GMC_UNROLL for (int v = 0; v < 4; v++) {
        const float4 av = pa[v], bv = pb[v];
        acc[4 * v + 0] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bv.x, acc[4 * v + 0], 0, 0, 0);
        acc[4 * v + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bv.y, acc[4 * v + 1], 0, 0, 0);
        acc[4 * v + 2] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bv.z, acc[4 * v + 2], 0, 0, 0);
        acc[4 * v + 3] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bv.w, acc[4 * v + 3], 0, 0, 0);
      }
    }
  }
  // the products back to the quad layout, 8 elements at a time: this lane's
  // entries t = 4k + q of every block (b, a) of its element
  GMC_UNROLL for (int h = 0; h < 2; h++) {
    wave_lds_sync();
    GMC_UNROLL for (int w = 0; w < 8; w++)
      GMC_UNROLL for (int r = 0; r < 4; r++)
        sw[w * 256 + (4 * (lane >> 4) + r) * 16 + (lane & 15)] = acc[8 * h + w][r];
    wave_lds_sync();
    if ((ew >> 3) == h) {
      const float* c = sw + (ew & 7) * 256;
      GMC_UNROLL for (int b = 0; b < NEN; b++)
        GMC_UNROLL for (int a2 = 0; a2 < NEN; a2++)
          GMC_UNROLL for (int k = 0; k < (NB + 3) / 4; k++) {
            const int t = 4 * k + q < NB ? 4 * k + q : 0, i = t / N, j = t - i * N;
            lsq[b][a2][k] = c[(b * N + i) * 16 + a2 * N + j];
          }
    }
  }

}

template <int DIM, int NC, bool RHS>
__global__ __launch_bounds__(256, 2) void k_interior_mfma(const AsmArgs a) {
  using S = Shape<DIM, NC>;
  using L = EbLayout<DIM, NC>;
  using TM = float;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE, NB = N * N;
  static_assert(DIM == 2 && NE <= 16 && NEN * N <= 16, "one 16 x 16 tile per element");
  __shared__ __align__(16) float st[4 * mfma::WAVE];
  const int q = threadIdx.x & 3;
  const int lane = threadIdx.x & 63;
  const int ew = lane >> 2;  // the quad's element within the wave
  float* sw = st + (threadIdx.x >> 6) * mfma::WAVE;
  const long long gid = (long long)blockIdx.x * (blockDim.x >> 2) + (threadIdx.x >> 2);
  const bool live = gid < a.n_items;
  const long long e = a.items[live ? gid : a.n_items - 1];
  int eslot[NEN * NEN];
  GMC_UNROLL for (int k = 0; k < NEN * NEN; k++) eslot[k] = a.slot[e * NEN * NEN + k];
  const DevParams& P = *a.params;
  Material<double> M;
  load_material<double, NC>(P, M);

  int en[NEN];
  double xe[NEN * DIM], Ie[NE], Pe[NE];
  gather<DIM, NC>(a, e, en, xe, Ie, Pe);
  Geo<DIM> g;
  const bool ok = geometry<DIM>(xe, g);
  if (!ok && live && q == 0) atomicOr(a.status, 1);
  const bool store = live && ok;
  Point<double, DIM, NC> pt;
  const int qp = q < S::NQ ? q : S::NQ - 1;
  const int stt = point_setup<double, DIM, NC>(P, M, g, qp, q < S::NQ ? 1.0 : 0.0, Ie, Pe, pt,
                                               magma_row<NC>(a.mt, e * S::NQ + qp, 2));
  if (stt && q < S::NQ && store) atomicOr(a.status, stt);

  unsigned dm[NEN];
  GMC_UNROLL for (int n = 0; n < NEN; n++) dm[n] = a.dmask[en[n]];

  if (RHS) {
    double r[NE];
    point_residual<double, DIM, NC>(P, pt, g, r);
    double* ebr = a.ebr + e * L::RS;
    double rs[(NE + 3) / 4];
    quad_rs<double, NE>(r, q, rs);
    GMC_UNROLL for (int k = 0; k < (NE + 3) / 4; k++) {
      const int t = 4 * k + q;
      if (t >= NE || !store) continue;
      const int b = t / N, jb = t - b * N;
      const bool blocked = (pick<NEN>(dm, b) >> jb) & 1u;
      ebr[t] = blocked ? 0.0 : -rs[k];
    }
  }

  MatPoint<TM, DIM, NC> mp;
  make_matpoint<TM, DIM, NC>(P, pt, g, mp);
  TM lsq[NEN][NEN][(NB + 3) / 4];
  lsq_mfma<DIM, NC>(mp, sw, lane, lsq);

  // Galerkin, diffusion and DC terms; stores as k_interior's element-buffer mode
  TM* ebm = reinterpret_cast<TM*>(a.ebm);
  TM tpart = TM(0.0);
  GMC_UNROLL for (int ia = 0; ia < NEN; ia++)
    GMC_UNROLL for (int ib = 0; ib < NEN; ib++) {
      TM m[N][N];
      point_block2_full<TM, DIM, NC, false>(mp, nullptr, nullptr, ib, ia, m);
      TM mv[(NB + 3) / 4];
      quad_rs<TM, NB>(&m[0][0], q, mv);
      GMC_UNROLL for (int k = 0; k < (NB + 3) / 4; k++) {
        const int t = 4 * k + q, jb = t / N, ja = t - jb * N;
        if (t >= NB) continue;
        const TM v = mv[k] + lsq[ib][ia][k];
        if (ib == ia && jb == ja) tpart += v;
        if (store) ebm[(long long)eslot[ib * NEN + ia] * NB + t] = ((dm[ib] >> jb) & 1u) ? TM(0) : v;
      }
    }
  const TM trace = TM(1.0) + quad_sum(tpart);
  // Dirichlet rows: diagonal = avg_val (scatter.hpp:65-78), overwriting the 0
  const TM avg = trace / TM(NE);
  GMC_UNROLL for (int ib = 0; ib < NEN; ib++) {
    if (!dm[ib] || !store) continue;
    GMC_UNROLL for (int jb = 0; jb < N; jb++) {
      const int t = jb * N + jb;
      if ((t & 3) == q && ((dm[ib] >> jb) & 1u)) ebm[(long long)eslot[ib * NEN + ib] * NB + t] = avg;
    }
  }
}

// ---------------------------------------------------------------------------
// boundary elements (execute_b + RM_b, fluid_2d.hpp:209-245, 341-364).  The
// (side, side point) pairs of an element are dealt round-robin to the 4 lanes;
// each round adds its quad-summed partial contribution.  Boundary integrals
// are evaluated in FP64 in every tier (they are ~0.1-2% of the elements) and
// stored in the element-matrix precision TM.
// ---------------------------------------------------------------------------
template <int DIM, int NC, bool MAT, bool RHS, int MODE, typename TM>
__global__ __launch_bounds__(256) void k_boundary(const AsmArgs a) {
  using S = Shape<DIM, NC>;
  using L = EbLayout<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE, NB = N * N;
  constexpr bool OUT = MODE == MODE_DENSE;
  constexpr bool TR = MODE == MODE_VJPT;
  constexpr bool JVP = MODE == MODE_JVP || TR;
  constexpr bool EBUF = MODE == MODE_EBUF || JVP;
  constexpr bool HALO = MODE == MODE_HALO;
  const int q = threadIdx.x & 3;
  const long long gid = (long long)blockIdx.x * (blockDim.x >> 2) + (threadIdx.x >> 2);
  if (gid >= a.n_items) return;
  const long long e = a.items[gid];
  const DevParams& P = *a.params;
  Material<double> M;
  load_material<double, NC>(P, M);
  int en[NEN];
  double xe[NEN * DIM], Ie[NE];
  gather<DIM, NC, false>(a, e, en, xe, Ie, nullptr);
  double xv[JVP ? NE : 1], yv[JVP ? NE : 1];
  if constexpr (JVP) {
    GMC_UNROLL for (int n = 0; n < NEN; n++)
      GMC_UNROLL for (int k = 0; k < N; k++) {
        xv[n * N + k] = a.xvec[(long long)en[n] * N + k];
        yv[n * N + k] = 0.0;
      }
  }
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) {
    if (q == 0) atomicOr(a.status, 1);
    return;
  }
  unsigned dm[NEN];
  GMC_UNROLL for (int n = 0; n < NEN; n++) dm[n] = OUT ? 0u : a.dmask[en[n]];
  const int s0 = a.side_start[e], s1 = a.side_start[e + 1];
  const int npairs = (s1 - s0) * S::NSQ;
  // matrix-free modes sum the trace in every lane; the storing modes keep
  // this lane's diagonal entries and quad-sum them at the end
  double trace = 1.0, tpart = 0.0;
  double* ebr = EBUF ? a.ebr + e * L::RS : nullptr;
  TM* ebm = EBUF ? reinterpret_cast<TM*>(a.ebm) : nullptr;
  const int* eslot = EBUF ? a.slot + e * NEN * NEN : nullptr;
  // halo slots of this boundary element: its blocks, then its nodes (each
  // written by this element alone: assigned at the first pass over its side
  // points, added to at the later ones)
  const int* hs = HALO ? a.hslot + gid * (NEN * NEN + NEN) : nullptr;
  for (int base = 0; base < npairs; base += 4) {
    const int t = base + q;
    const bool act = t < npairs;
    const int tt = act ? t : (npairs - 1);
    const int sidx = s0 + tt / S::NSQ;
    BPoint<double, DIM, NC> bp;
    bpoint_setup<double, DIM, NC>(P, M, g, xe, a.side_local[sidx], tt % S::NSQ, a.side_flag[sidx],
                                  act ? 1.0 : 0.0, Ie, bp,
                                  magma_row<NC>(a.mtb, (long long)sidx * S::NSQ + tt % S::NSQ, 1));
    if (RHS) {
      double r[NE];
      bpoint_residual<double, DIM, NC>(bp, r);
      double rs[(NE + 3) / 4];
      quad_rs<double, NE>(r, q, rs);  // rs[k]: entry 4k + q
      GMC_UNROLL for (int k = 0; k < (NE + 3) / 4; k++) {
        const int ti = 4 * k + q;
        if (ti >= NE) continue;
        const int b = ti / N, j = ti - b * N;
        const double rr = rs[k];
        const bool blocked = (pick<NEN>(dm, b) >> j) & 1u;
        if (OUT) a.out_r[gid * NE + ti] -= rr;
        else if (EBUF) ebr[ti] -= blocked ? 0.0 : rr;
        else if (HALO) {
          double* dst = a.halo_r + (long long)hs[NEN * NEN + b] * N + j;
          const double v = blocked ? 0.0 : -rr;
          *dst = base ? *dst + v : v;
        }
        else if (!blocked) a.rhs[(long long)pick<NEN>(en, b) * N + j] -= rr;
      }
    }
    if (MAT) {
      // column by column: the NEN blocks of column ia are N_b C_ia
      GMC_UNROLL for (int ia = 0; ia < NEN; ia++) {
        double C[N][N];
        bpoint_col<double, DIM, NC>(bp, g, ia, C);
        GMC_UNROLL for (int ib = 0; ib < NEN; ib++) {
          double m[N][N];
          GMC_UNROLL for (int jb = 0; jb < N; jb++)
            GMC_UNROLL for (int ja = 0; ja < N; ja++) m[jb][ja] = bp.sh[ib] * C[jb][ja];
          if constexpr (JVP) {
            GMC_UNROLL for (int jb = 0; jb < N; jb++)
              GMC_UNROLL for (int ja = 0; ja < N; ja++) {
                const double mm = quad_sum(m[jb][ja]);
                if (ib == ia && jb == ja) trace += mm;
                if constexpr (TR) {
                  if (!((dm[ib] >> jb) & 1u)) yv[ia * N + ja] += mm * xv[ib * N + jb];
                } else {
                  yv[ib * N + jb] += mm * xv[ia * N + ja];
                }
              }
          } else {
            // reduce-scatter: this lane's entries 4k + q of the block
            const long long blk = (OUT || EBUF || HALO) ? 0 : a.blkmap[e * NEN * NEN + ib * NEN + ia];
            double mv[(NB + 3) / 4];
            quad_rs<double, NB>(&m[0][0], q, mv);
            GMC_UNROLL for (int k = 0; k < (NB + 3) / 4; k++) {
              const int ti = 4 * k + q, jb = ti / N, ja = ti - jb * N;
              if (ti >= NB) continue;
              const double mm = mv[k];
              if (ib == ia && jb == ja) tpart += mm;
              const bool blocked = (dm[ib] >> jb) & 1u;
              if (OUT) a.out_m[gid * NE * NE + (long long)(ib * N + jb) * NE + ia * N + ja] += mm;
              else if (EBUF) {
                if (!blocked) ebm[(long long)eslot[ib * NEN + ia] * NB + ti] += TM(mm);
              } else if (HALO) {
                double* dst = a.halo_m + (long long)hs[ib * NEN + ia] * NB + ti;
                const double v = blocked ? 0.0 : mm;
                *dst = base ? *dst + v : v;
              } else if (!blocked) a.vals[blk * NB + ti] += mm;
            }
          }
        }
      }
    }
  }
  if constexpr (JVP) {
    const double avg = trace / NE;
    GMC_UNROLL for (int t = 0; t < NE; t++)
      if ((t & 3) == q) {
        const bool blocked = (dm[t / N] >> (t % N)) & 1u;
        if (TR) ebr[t] += yv[t] + (blocked ? avg * xv[t] : 0.0);
        else ebr[t] += blocked ? avg * xv[t] : yv[t];
      }
  } else if (MAT && !OUT) {
    trace += quad_sum(tpart);
    const double avg = trace / NE;
    GMC_UNROLL for (int ib = 0; ib < NEN; ib++) {
      if (!dm[ib]) continue;
      GMC_UNROLL for (int jb = 0; jb < N; jb++) {
        const int t = jb * N + jb;
        if ((t & 3) == q && ((dm[ib] >> jb) & 1u)) {
          if (EBUF) ebm[(long long)eslot[ib * NEN + ib] * NB + t] += TM(avg);
          else if (HALO) a.halo_m[(long long)hs[ib * NEN + ib] * NB + t] += avg;
          else a.vals[(long long)a.blkmap[e * NEN * NEN + ib * NEN + ib] * NB + t] += avg;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// patch-local assembly: one workgroup = the pe elements (4 lanes each) of a
// compact patch.  The element work is that of k_interior; its blocks and rhs
// parts are summed into the workgroup's LDS slots (FP64 in every tier) one
// local block index at a time, phase by phase, so every slot adds its
// contributions in a fixed order; a workgroup barrier separates the phases.
// The finished slots go straight to the BSR values / rhs when the patch owns
// them, else to halo slots for k_halo_gather.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void patch_lds_fence() { __syncthreads(); }

// *p += v in LDS.  2D: one ds_add_f64 (no return), so that a phase costs one
// LDS round trip, not a read, its wait and a write; within a phase no two
// lanes add to the same entry and the phases are ordered by the barriers, so
// the sums are those of the read-modify-write, bit for bit.  3D keeps the
// read-modify-write.
template <int DIM>
__device__ __forceinline__ void patch_lds_add(double* p, double v) {
  if constexpr (DIM == 2)
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else
    *p += v;
}

// The store loop of a patch: acc holds n slots of NV values; slot sl goes to
// out[dst[sl] NV ..] (dst >= 0) or halo[(-dst[sl] - 1) NV ..].  Lane
// (sl0, j) = (lane / NV, lane % NV) stores entry j of slots sl0, sl0 + P, ...
// with P = nt / NV slots per pass (lanes past P NV idle).
template <int NV>
__device__ __forceinline__ void patch_store(const double* acc, const int* dst, int n, int lane, int nt,
                                            double* out, double* halo) {
  const int j = lane % NV, per = nt / NV;
  for (int sl = lane < per * NV ? lane / NV : n; sl < n; sl += per) {
    const int d = dst[sl];
    const double v = acc[sl * NV + j];
    if (d >= 0) out[(long long)d * NV + j] = v;
    else halo[(long long)(-d - 1) * NV + j] = v;
  }
}

// The element data of one patch lane (4 lanes per element): everything the
// point physics reads from global memory, loaded one patch ahead.
template <int DIM, int NC>
struct PatchElem {
  using S = Shape<DIM, NC>;
  long long e;
  bool valid;
  int en[S::NEN];
  double xe[S::NEN * DIM], Ie[S::NE], Pe[S::NE];
  unsigned ec[S::NEN * S::NEN], nc[S::NEN], dm[S::NEN];
};

template <int DIM, int NC>
__device__ __forceinline__ void load_patch_elem(const AsmArgs& a, const PatchArgs& pa, int p,
                                                int lane, PatchElem<DIM, NC>& d) {
  using S = Shape<DIM, NC>;
  constexpr int NEN = S::NEN, NL = NEN * NEN;
  const long long gid = (long long)p * pa.pe + (lane >> 2);
  d.valid = gid < a.n_items;
  const long long gsafe = d.valid ? gid : a.n_items - 1;
  d.e = a.items[gsafe];
  GMC_UNROLL for (int k = 0; k < NL; k++) d.ec[k] = pa.ecode[gsafe * NL + k];
  GMC_UNROLL for (int k = 0; k < NEN; k++) d.nc[k] = pa.ncode[gsafe * NEN + k];
  gather<DIM, NC>(a, d.e, d.en, d.xe, d.Ie, d.Pe);
  GMC_UNROLL for (int n = 0; n < NEN; n++) d.dm[n] = a.dmask[d.en[n]];
}

// Workgroup b assembles the patches patch_seq gives it (G = gridDim.x).  The
// kernels built for one wave per SIMD (FP64 matrix) run a persistent grid (G
// = the workgroups that fit on the device): with one wave per SIMD nothing
// else hides the latency of a new patch's dependent loads (item ->
// connectivity -> coordinates and state), so each workgroup issues the next
// patch's loads after its matrix phase and before its stores, when its
// registers are free.  The two-wave kernels launch one workgroup per patch
// (G = the patch count): the other workgroup on the CU hides the loads, and
// carrying the next patch's data costs them registers.  One workgroup is
// 4 * pa.pe lanes (at most 256).
template <int DIM, bool MAT, typename TM>
constexpr bool patch_prefetch() {
  return waves_per_simd<DIM, MAT, TM>() == 1;
}

// The patches of workgroup b, first, first + step, ... below end.  With xcd:
// consecutive patches (neighbours in the Morton order, sharing nodes and
// blocks) on one XCD.  Workgroups are dealt to the 8 XCDs round-robin by
// index, so XCD x = b mod 8 takes the contiguous patch range [lo_x, hi_x) and
// its workgroups stride through it; used when every XCD holds as many
// workgroups as patches (G = P) or the same number (G mod 8 = 0).  Otherwise
// (and without xcd) b, b + G, b + 2G, ...
struct PatchSeq {
  int first, step, end;
};
__device__ __forceinline__ PatchSeq patch_seq(int b, int G, int P, bool xcd) {
  if (!xcd || !(G == P || (G & 7) == 0)) return {b, G, P};
  const int x = b & 7, w = b >> 3, q = P >> 3, r = P & 7;
  const int lo = x * q + (x < r ? x : r), hi = lo + q + (x < r ? 1 : 0);
  const int nw = (G >> 3) + (x < (G & 7) ? 1 : 0);
  return {lo + w, nw, hi};
}

// MF: the FP32 tier's least-squares term on the matrix cores (lsq_mfma), its
// staging areas in the block-slot LDS before the slots are cleared
template <int DIM, int NC, bool MAT, bool RHS, typename TM, bool MF = false>
__global__ __launch_bounds__(256, (waves_per_simd<DIM, MAT, TM>())) void k_patch(const AsmArgs a, const PatchArgs pa) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE, NB = N * N, NL = NEN * NEN;
  static_assert(!MF || (MAT && sizeof(TM) == 4), "matrix-core least squares: FP32 tier only");
  extern __shared__ __align__(16) double lds[];
  const int lane = threadIdx.x;
  const int q = lane & 3;
  const int nt = blockDim.x;
  double* accm = lds;
  double* accr = lds + (MF ? patch_acc_doubles(NB, pa.max_slots, nt) : (size_t)pa.max_slots * NB);
  // slot destinations, staged in LDS at the start so the store loop at the
  // end does not wait on their loads
  int* dstm = reinterpret_cast<int*>(accr + (size_t)pa.max_nodes * N);
  int* dstr = dstm + pa.max_slots;
  const DevParams& P = *a.params;
  constexpr bool PF = patch_prefetch<DIM, MAT, TM>();
  PatchElem<DIM, NC> cur;
  // XCD-contiguous patches for the one-patch-per-workgroup 2D kernels (FP32
  // tier, config 1: 3.79 -> 3.74 ms per assembly); the persistent FP64 kernel
  // measured the same either way and keeps b, b + G, ...
  const PatchSeq seq = patch_seq((int)blockIdx.x, (int)gridDim.x, pa.n_patches, DIM == 2 && !PF);
  if (PF && seq.first < seq.end) load_patch_elem<DIM, NC>(a, pa, seq.first, lane, cur);
  for (int p = seq.first; p < seq.end; p += seq.step) {
    if (!PF) load_patch_elem<DIM, NC>(a, pa, p, lane, cur);
    const int s0 = pa.slot_ptr[p], ns = pa.slot_ptr[p + 1] - s0;
    const int n0 = pa.node_ptr[p], nn = pa.node_ptr[p + 1] - n0;
    // the previous patch's store loop has read its LDS slots
    if (p != seq.first) patch_lds_fence();
    if (MAT) {
      if (!MF)
        for (int i = lane; i < ns * NB; i += nt) accm[i] = 0.0;
      for (int i = lane; i < ns; i += nt) dstm[i] = pa.slot_dest[s0 + i];
    }
    if (RHS) {
      for (int i = lane; i < nn * N; i += nt) accr[i] = 0.0;
      for (int i = lane; i < nn; i += nt) dstr[i] = pa.node_dest[n0 + i];
    }
    // material constants reloaded per patch (uniform, cached) rather than
    // held in registers across the matrix phase
    Material<double> M;
    load_material<double, NC>(P, M);
    const bool valid = cur.valid;
    const unsigned* ec = cur.ec;
    const unsigned* nc = cur.nc;
    const unsigned* dm = cur.dm;
    const double* xe = cur.xe;
    const double* Ie = cur.Ie;
    const double* Pe = cur.Pe;
    Geo<DIM> g;
    bool act = valid;
    if (!geometry<DIM>(xe, g)) {
      if (valid && q == 0) atomicOr(a.status, 1);
      act = false;
    }
    Point<double, DIM, NC> pt;
    const int qp = q < S::NQ ? q : S::NQ - 1;
    // the 3D patch kernels (not the 3D default scatter) keep tau's inverse by
    // the pivoted LU: built with the closed form, a run of their parity test
    // faulted the GPU once, with no cause found in the code; with the LU they
    // are the kernels that passed that test before
    const int st = point_setup<double, DIM, NC, DIM == 3>(P, M, g, qp, q < S::NQ ? 1.0 : 0.0, Ie, Pe, pt,
                                                          magma_row<NC>(a.mt, cur.e * S::NQ + qp, 2));
    if (valid && st && q < S::NQ) atomicOr(a.status, st);
    patch_lds_fence();  // LDS cleared

    if (RHS) {
      double r[NE];
      point_residual<double, DIM, NC>(P, pt, g, r);
      double rv[(NE + 3) / 4];
      quad_rs<double, NE>(r, q, rv);  // rv[k]: entry 4k + q
      GMC_UNROLL for (int k = 0; k < (NE + 3) / 4; k++) {
        const int t = 4 * k + q;
        // dm[t / N] by selects: a runtime index would put the mask in scratch
        unsigned dmt = dm[0];
        GMC_UNROLL for (int bb = 1; bb < NEN; bb++) dmt = (t / N == bb) ? dm[bb] : dmt;
        rv[k] = (t < NE && ((dmt >> (t % N)) & 1u)) ? 0.0 : -rv[k];
      }
      // one local node index b at a time (the plan's phases are ranks among the
      // elements that add into the same node slot at the same b); lanes of
      // other waves may add to the same slot at another b, so every step ends
      // at a workgroup barrier
      GMC_UNROLL for (int b = 0; b < NEN; b++) {
        const int ln = (int)(nc[b] & 0xFFFu), myph = (int)(nc[b] >> 12);
        for (int ph = 0; ph <= pa.max_nphase; ph++) {
          if (b || ph) patch_lds_fence();
          if (act && myph == ph) {
            GMC_UNROLL for (int k = 0; k < (NE + 3) / 4; k++) {
              const int t = 4 * k + q;
              if (t < NE && t / N == b) patch_lds_add<DIM>(&accr[ln * N + t % N], rv[k]);
            }
          }
        }
      }
    }

    if (MAT) {
      MatPoint<TM, DIM, NC> mp;
      make_matpoint<TM, DIM, NC>(P, pt, g, mp);
      // this lane's part of the element trace (diagonal entries 4k + q of the
      // diagonal blocks), quad-summed after the blocks
      TM tpart = TM(0.0);
      MatPhase<TM, DIM, NC> ph2;
      TM lsq[MF ? NEN : 1][MF ? NEN : 1][(NB + 3) / 4];
      if constexpr (MF) {
        lsq_mfma<DIM, NC>(mp, reinterpret_cast<float*>(lds) + (lane >> 6) * mfma::WAVE, lane & 63, lsq);
        patch_lds_fence();  // every wave is done with its staging area
        for (int i = lane; i < ns * NB; i += nt) accm[i] = 0.0;
        patch_lds_fence();
      } else {
        ph2.init(mp);
      }
      GMC_UNROLL for (int ia = 0; ia < NEN; ia++) {
        if constexpr (!MF) ph2.col(mp, ia);
        GMC_UNROLL for (int ib = 0; ib < NEN; ib++) {
          TM m[N][N];
          if constexpr (MF) point_block2_full<TM, DIM, NC, false>(mp, nullptr, nullptr, ib, ia, m);
          else ph2.block(mp, ib, ia, m);
          TM sv[(NB + 3) / 4];
          quad_rs<TM, NB>(&m[0][0], q, sv);  // sv[k]: entry t = 4k + q of the block
          if constexpr (MF)
            GMC_UNROLL for (int k = 0; k < (NB + 3) / 4; k++) sv[k] += lsq[MF ? ib : 0][MF ? ia : 0][k];
          double mv[(NB + 3) / 4];
          GMC_UNROLL for (int k = 0; k < (NB + 3) / 4; k++) {
            const int t = 4 * k + q, jb = t / N, ja = t - jb * N;
            if (ib == ia && t < NB && jb == ja) tpart += sv[k];
            mv[k] = ((dm[ib] >> jb) & 1u) ? 0.0 : (double)sv[k];
          }
          const unsigned code = ec[ib * NEN + ia];
          const int ls = (int)(code & 0xFFFu), myph = (int)(code >> 12);
          for (int ph = 0; ph <= pa.max_phase; ph++) {
            patch_lds_fence();
            if (act && myph == ph) {
              GMC_UNROLL for (int k = 0; k < (NB + 3) / 4; k++) {
                const int t = 4 * k + q;
                if (t < NB) patch_lds_add<DIM>(&accm[ls * NB + t], mv[k]);
              }
            }
          }
        }
      }
      // Dirichlet rows: avg_val on the blocked diagonals (scatter.hpp:65-78)
      const TM trace = TM(1.0) + quad_sum(tpart);
      const double avg = (double)(trace / TM(NE));
      GMC_UNROLL for (int ib = 0; ib < NEN; ib++) {
        const unsigned code = ec[ib * NEN + ib];
        const int ls = (int)(code & 0xFFFu), myph = (int)(code >> 12);
        for (int ph = 0; ph <= pa.max_phase; ph++) {
          patch_lds_fence();
          if (act && myph == ph && dm[ib]) {
            GMC_UNROLL for (int jb = 0; jb < N; jb++) {
              const int t = jb * N + jb;
              if ((t & 3) == q && ((dm[ib] >> jb) & 1u)) patch_lds_add<DIM>(&accm[ls * NB + t], avg);
            }
          }
        }
      }
    }
    // the next patch's element data, in flight during the stores (the last
    // iteration reloads its own patch, so that cur is always replaced and
    // none of it stays live across the body)
    const int pn = p + seq.step;
    PatchElem<DIM, NC> nxt;
    if (PF) load_patch_elem<DIM, NC>(a, pa, pn < seq.end ? pn : p, lane, nxt);
    patch_lds_fence();
    if constexpr (DIM == 2) {
      // lane = (slot of the pass, entry): each pass stores nt / NB whole
      // slots, so the loop carries no division and one destination per slot
      if (MAT) patch_store<NB>(accm, dstm, ns, lane, nt, a.vals, pa.halo_m);
      if (RHS) patch_store<N>(accr, dstr, nn, lane, nt, a.rhs, pa.halo_r);
    } else {
    if (MAT)
      for (int i = lane; i < ns * NB; i += nt) {
        const int sl = i / NB, j = i - sl * NB;
        const int d = dstm[sl];
        if (d >= 0) a.vals[(long long)d * NB + j] = accm[i];
        else pa.halo_m[(long long)(-d - 1) * NB + j] = accm[i];
      }
    if (RHS)
      for (int i = lane; i < nn * N; i += nt) {
        const int sl = i / N, j = i - sl * N;
        const int d = dstr[sl];
        if (d >= 0) a.rhs[(long long)d * N + j] = accr[i];
        else pa.halo_r[(long long)(-d - 1) * N + j] = accr[i];
      }
    }
    if (!PF) break;  // one patch per workgroup: no loop to carry values across
    cur = nxt;
  }
}

template <int DIM, int NC, typename TM>
hipError_t launch_shape(int which, int flags, const AsmArgs& a, hipStream_t s) {
  if (a.n_items == 0) return hipSuccess;
  const bool mat = flags & GMC_ASM_MATRIX, rhs = flags & GMC_ASM_RHS;
  const bool bd = which & ASM_BOUNDARY;
  // elements (quads) per block: 64 (256 threads); boundary launches cover a
  // few thousand elements each, so they take 16 per block to spread over
  // more CUs
  const int per_block = bd ? 16 : 64;
  const dim3 grid((unsigned)((a.n_items + per_block - 1) / per_block)), block(4 * per_block);
  const int mode = (which & ASM_OUT) ? MODE_DENSE
                   : (which & ASM_JVP) ? MODE_JVP
                   : (which & ASM_VJPT) ? MODE_VJPT
                   : ((which & ASM_EBUF) ? MODE_EBUF : MODE_SCATTER);
#define GMC_LAUNCH(KERN, M_, R_, O_) \
  hipLaunchKernelGGL((KERN<DIM, NC, M_, R_, O_, TM>), grid, block, 0, s, a)
#define GMC_MODES(KERN)                                                       \
  if (mode == MODE_DENSE) GMC_LAUNCH(KERN, true, true, MODE_DENSE);          \
  else if (mode == MODE_JVP) GMC_LAUNCH(KERN, true, false, MODE_JVP);        \
  else if (mode == MODE_VJPT) GMC_LAUNCH(KERN, true, false, MODE_VJPT);      \
  else if (mode == MODE_EBUF) {                                               \
    if (mat && rhs) GMC_LAUNCH(KERN, true, true, MODE_EBUF);                  \
    else if (mat) GMC_LAUNCH(KERN, true, false, MODE_EBUF);                   \
    else GMC_LAUNCH(KERN, false, true, MODE_EBUF);                            \
  } else {                                                                    \
    if (mat && rhs) GMC_LAUNCH(KERN, true, true, MODE_SCATTER);               \
    else if (mat) GMC_LAUNCH(KERN, true, false, MODE_SCATTER);                \
    else GMC_LAUNCH(KERN, false, true, MODE_SCATTER);                         \
  }
  if constexpr (DIM == 2 && NC == 2 && sizeof(TM) == 4) {
    if (!bd && (which & ASM_MFMA) && mode == MODE_EBUF && mat) {
      if (rhs) hipLaunchKernelGGL((k_interior_mfma<DIM, NC, true>), grid, block, 0, s, a);
      else hipLaunchKernelGGL((k_interior_mfma<DIM, NC, false>), grid, block, 0, s, a);
      return hipGetLastError();
    }
  }
  if (!bd) {
    GMC_MODES(k_interior)
  } else if (which & ASM_HALO) {
    // the patch scatter's boundary elements, into their own halo slots
    if (mat && rhs) GMC_LAUNCH(k_boundary, true, true, MODE_HALO);
    else if (mat) GMC_LAUNCH(k_boundary, true, false, MODE_HALO);
    else GMC_LAUNCH(k_boundary, false, true, MODE_HALO);
  } else {
    GMC_MODES(k_boundary)
  }
#undef GMC_MODES
#undef GMC_LAUNCH
  return hipGetLastError();
}

template <int DIM, int NC, bool MAT, bool RHS, typename TM, bool MF = false>
hipError_t launch_patch_k(const AsmArgs& a, const PatchArgs& pa, hipStream_t s) {
  constexpr int N = Shape<DIM, NC>::NDOF;
  const size_t lds = patch_lds_bytes(N, pa, MF);
  auto kern = k_patch<DIM, NC, MAT, RHS, TM, MF>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  // persistent grid for the prefetching kernels: the workgroups that fit on
  // the device at once (at most one per patch)
  int dev = 0, n_cu = 0, per_cu = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern),
                                                        4 * pa.pe, lds)) != hipSuccess)
    return e;
  const long long fit = (long long)n_cu * (per_cu > 0 ? per_cu : 1);
  const unsigned grid = (unsigned)(patch_prefetch<DIM, MAT, TM>() && fit < pa.n_patches ? fit
                                                                                         : pa.n_patches);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(4 * pa.pe), lds, s, a, pa);
  return hipGetLastError();
}

template <int DIM, int NC, typename TM>
hipError_t launch_patch_shape(int flags, const AsmArgs& a, const PatchArgs& pa,
                                     hipStream_t s) {
  if (pa.n_patches == 0) return hipSuccess;
  const bool mat = flags & GMC_ASM_MATRIX, rhs = flags & GMC_ASM_RHS;
  if constexpr (DIM == 2 && NC == 2 && sizeof(TM) == 4) {
    if (mat && (flags & PATCH_MFMA)) {
      if (rhs) return launch_patch_k<DIM, NC, true, true, TM, true>(a, pa, s);
      return launch_patch_k<DIM, NC, true, false, TM, true>(a, pa, s);
    }
  }
  if (mat && rhs) return launch_patch_k<DIM, NC, true, true, TM>(a, pa, s);
  if (mat) return launch_patch_k<DIM, NC, true, false, TM>(a, pa, s);
  return launch_patch_k<DIM, NC, false, true, TM>(a, pa, s);
}

}  // namespace gmc
