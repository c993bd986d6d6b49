// gmc_assemble.hip -- FP64 element assembly kernels for gfx950.
//
// One element per 4-lane group ("quad"); lane q of the quad evaluates
// interior quadrature point q (fluid_2d.hpp:177-199 loops the points of one
// element serially).  Triangles have 3 points, so lane 3 duplicates point 2
// with weight 0; tetrahedra use all 4 lanes.  Each lane holds its point's
// coefficient matrices (DUDY, A_k, tau) in registers and produces one
// ndof x ndof block of the element matrix at a time; the quad sums the block
// with DPP quad permutes and the four lanes write disjoint entries of it.
//
// Scatter (scatter.hpp:37-98) is deterministic without atomics: elements are
// launched one colour at a time, and no two elements of a colour share a node,
// so each BSR block entry has one writer per launch and launches are ordered on
// the stream.  The Dirichlet rule (avg_val = (1 + sum_i m_ii)/n on blocked
// diagonals, zero rows, zero rhs) is applied per element and per integral
// (interior, boundary), as the reference applies it per scatter call.
#include <hip/hip_runtime.h>

#include "gmc_element.hpp"
#include "gmc_kernels.hpp"

namespace gmc {

// quad_perm DPP: lane i reads lane i^1 (0xB1) or lane i^2 (0x4E) of its quad
__device__ __forceinline__ double dpp_quad(double x, int ctrl_is_xor2) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  if (ctrl_is_xor2) {
    lo = __builtin_amdgcn_mov_dpp(lo, 0x4E, 0xF, 0xF, false);
    hi = __builtin_amdgcn_mov_dpp(hi, 0x4E, 0xF, 0xF, false);
  } else {
    lo = __builtin_amdgcn_mov_dpp(lo, 0xB1, 0xF, 0xF, false);
    hi = __builtin_amdgcn_mov_dpp(hi, 0xB1, 0xF, 0xF, false);
  }
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double quad_sum(double x) {
  x += dpp_quad(x, 0);
  x += dpp_quad(x, 1);
  return x;
}

// model::extract_element_dofs (model.hpp:116-132): node ids, coordinates and
// the element vectors of the I (and P) state.
template <int DIM, int NC>
__device__ __forceinline__ void gather(const AsmArgs& a, long long e, int* en, double* xe,
                                       double* Ie, double* Pe) {
  using S = Shape<DIM, NC>;
  GMC_UNROLL for (int n = 0; n < S::NEN; n++) {
    en[n] = a.conn[e * S::NEN + n];
    GMC_UNROLL for (int k = 0; k < DIM; k++) xe[n * DIM + k] = a.coords[(long long)en[n] * DIM + k];
    GMC_UNROLL for (int k = 0; k < S::NDOF; k++) {
      Ie[n * S::NDOF + k] = a.I[(long long)en[n] * S::NDOF + k];
      if (Pe) Pe[n * S::NDOF + k] = a.P[(long long)en[n] * S::NDOF + k];
    }
  }
}

// ---------------------------------------------------------------------------
// interior elements of one colour
// ---------------------------------------------------------------------------
template <int DIM, int NC, bool MAT, bool RHS, bool OUT>
__global__ __launch_bounds__(256) void k_interior(const AsmArgs a) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE;
  const int q = threadIdx.x & 3;
  const long long gid = (long long)blockIdx.x * (blockDim.x >> 2) + (threadIdx.x >> 2);
  if (gid >= a.n_items) return;  // whole quads leave together
  const long long e = a.items[gid];
  const DevParams& P = *a.params;
  Material<double> M;
  load_material<double, NC>(P, M);

  int en[NEN];
  double xe[NEN * DIM], Ie[NE], Pe[NE];
  gather<DIM, NC>(a, e, en, xe, Ie, Pe);
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) {
    if (q == 0) atomicOr(a.status, 1);
    return;
  }
  Point<double, DIM, NC> pt;
  const int st = point_setup<double, DIM, NC>(P, M, g, q < S::NQ ? q : S::NQ - 1,
                                              q < S::NQ ? 1.0 : 0.0, Ie, Pe, pt);
  if (st && q < S::NQ) atomicOr(a.status, st);

  unsigned dm[NEN];
  GMC_UNROLL for (int n = 0; n < NEN; n++) dm[n] = OUT ? 0u : a.dmask[en[n]];

  if (RHS) {
    double r[NE];
    point_residual<double, DIM, NC>(P, pt, g, r);
    GMC_UNROLL for (int t = 0; t < NE; t++) {
      const double rr = quad_sum(r[t]);
      const int b = t / N, jb = t % N;
      if ((t & 3) == q) {
        if (OUT) a.out_r[gid * NE + t] = -rr;
        else if (!((dm[b] >> jb) & 1u)) a.rhs[(long long)en[b] * N + jb] -= rr;
      }
    }
  }

  if (MAT) {
    double trace = 1.0;
    GMC_UNROLL for (int ia = 0; ia < NEN; ia++) {
      double Ta[N][N];
      point_Ta<double, DIM, NC>(P, pt, g, ia, Ta);
      GMC_UNROLL for (int ib = 0; ib < NEN; ib++) {
        double m[N][N];
        point_block<double, DIM, NC>(P, pt, g, ib, ia, Ta, m);
        const long long blk = OUT ? 0 : a.blkmap[e * NEN * NEN + ib * NEN + ia];
        GMC_UNROLL for (int jb = 0; jb < N; jb++)
          GMC_UNROLL for (int ja = 0; ja < N; ja++) {
            const double mm = quad_sum(m[jb][ja]);
            const int t = jb * N + ja;
            if (ib == ia && jb == ja) trace += mm;
            if ((t & 3) == q) {
              if (OUT) a.out_m[gid * NE * NE + (long long)(ib * N + jb) * NE + ia * N + ja] = mm;
              else if (!((dm[ib] >> jb) & 1u)) a.vals[blk * N * N + t] += mm;
            }
          }
      }
    }
    if (!OUT) {
      // Dirichlet rows: diagonal = avg_val (scatter.hpp:65-78)
      const double avg = trace / NE;
      GMC_UNROLL for (int ib = 0; ib < NEN; ib++) {
        if (!dm[ib]) continue;
        const long long blk = a.blkmap[e * NEN * NEN + ib * NEN + ib];
        GMC_UNROLL for (int jb = 0; jb < N; jb++) {
          const int t = jb * N + jb;
          if ((t & 3) == q && ((dm[ib] >> jb) & 1u)) a.vals[blk * N * N + t] += avg;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// boundary elements of one colour (execute_b + RM_b).  The (side, side point)
// pairs of an element are dealt round-robin to the 4 lanes; each round adds
// its quad-summed partial contribution.
// ---------------------------------------------------------------------------
template <int DIM, int NC, bool MAT, bool RHS, bool OUT>
__global__ __launch_bounds__(256) void k_boundary(const AsmArgs a) {
  using S = Shape<DIM, NC>;
  constexpr int N = S::NDOF, NEN = S::NEN, NE = S::NE;
  const int q = threadIdx.x & 3;
  const long long gid = (long long)blockIdx.x * (blockDim.x >> 2) + (threadIdx.x >> 2);
  if (gid >= a.n_items) return;
  const long long e = a.items[gid];
  const DevParams& P = *a.params;
  Material<double> M;
  load_material<double, NC>(P, M);
  int en[NEN];
  double xe[NEN * DIM], Ie[NE];
  gather<DIM, NC>(a, e, en, xe, Ie, nullptr);
  Geo<DIM> g;
  if (!geometry<DIM>(xe, g)) {
    if (q == 0) atomicOr(a.status, 1);
    return;
  }
  unsigned dm[NEN];
  GMC_UNROLL for (int n = 0; n < NEN; n++) dm[n] = OUT ? 0u : a.dmask[en[n]];
  const int s0 = a.side_start[e], s1 = a.side_start[e + 1];
  const int npairs = (s1 - s0) * S::NSQ;
  double trace = 1.0;
  for (int base = 0; base < npairs; base += 4) {
    const int t = base + q;
    const bool act = t < npairs;
    const int tt = act ? t : (npairs - 1);
    const int sidx = s0 + tt / S::NSQ;
    BPoint<double, DIM, NC> bp;
    bpoint_setup<double, DIM, NC>(P, M, g, xe, a.side_local[sidx], tt % S::NSQ, a.side_flag[sidx],
                                  act ? 1.0 : 0.0, Ie, bp);
    if (RHS) {
      double r[NE];
      bpoint_residual<double, DIM, NC>(bp, r);
      GMC_UNROLL for (int ti = 0; ti < NE; ti++) {
        const double rr = quad_sum(r[ti]);
        const int b = ti / N, j = ti % N;
        if ((ti & 3) == q) {
          if (OUT) a.out_r[gid * NE + ti] -= rr;
          else if (!((dm[b] >> j) & 1u)) a.rhs[(long long)en[b] * N + j] -= rr;
        }
      }
    }
    if (MAT) {
      GMC_UNROLL for (int ib = 0; ib < NEN; ib++)
        GMC_UNROLL for (int ia = 0; ia < NEN; ia++) {
          double m[N][N];
          bpoint_block<double, DIM, NC>(bp, g, ib, ia, m);
          const long long blk = OUT ? 0 : a.blkmap[e * NEN * NEN + ib * NEN + ia];
          GMC_UNROLL for (int jb = 0; jb < N; jb++)
            GMC_UNROLL for (int ja = 0; ja < N; ja++) {
              const double mm = quad_sum(m[jb][ja]);
              const int ti = jb * N + ja;
              if (ib == ia && jb == ja) trace += mm;
              if ((ti & 3) == q) {
                if (OUT) a.out_m[gid * NE * NE + (long long)(ib * N + jb) * NE + ia * N + ja] += mm;
                else if (!((dm[ib] >> jb) & 1u)) a.vals[blk * N * N + ti] += mm;
              }
            }
        }
    }
  }
  if (MAT && !OUT) {
    const double avg = trace / NE;
    GMC_UNROLL for (int ib = 0; ib < NEN; ib++) {
      if (!dm[ib]) continue;
      const long long blk = a.blkmap[e * NEN * NEN + ib * NEN + ib];
      GMC_UNROLL for (int jb = 0; jb < N; jb++) {
        const int t = jb * N + jb;
        if ((t & 3) == q && ((dm[ib] >> jb) & 1u)) a.vals[blk * N * N + t] += avg;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int DIM, int NC>
static hipError_t launch_shape(int which, int flags, const AsmArgs& a, hipStream_t s) {
  if (a.n_items == 0) return hipSuccess;
  const int per_block = 64;  // elements (quads) per 256-thread block
  const dim3 grid((unsigned)((a.n_items + per_block - 1) / per_block)), block(256);
  const bool mat = flags & GMC_ASM_MATRIX, rhs = flags & GMC_ASM_RHS;
  const bool out = which & ASM_OUT;
  const bool bd = which & ASM_BOUNDARY;
#define GMC_LAUNCH(KERN, M_, R_, O_) hipLaunchKernelGGL((KERN<DIM, NC, M_, R_, O_>), grid, block, 0, s, a)
  if (!bd) {
    if (out) GMC_LAUNCH(k_interior, true, true, true);
    else if (mat && rhs) GMC_LAUNCH(k_interior, true, true, false);
    else if (mat) GMC_LAUNCH(k_interior, true, false, false);
    else GMC_LAUNCH(k_interior, false, true, false);
  } else {
    if (out) GMC_LAUNCH(k_boundary, true, true, true);
    else if (mat && rhs) GMC_LAUNCH(k_boundary, true, true, false);
    else if (mat) GMC_LAUNCH(k_boundary, true, false, false);
    else GMC_LAUNCH(k_boundary, false, true, false);
  }
#undef GMC_LAUNCH
  return hipGetLastError();
}

hipError_t launch_assemble(int dim, int nc, int which, int flags, const AsmArgs& a,
                           hipStream_t s) {
  if (dim == 2 && nc == 2) return launch_shape<2, 2>(which, flags, a, s);
  if (dim == 2 && nc == 3) return launch_shape<2, 3>(which, flags, a, s);
  if (dim == 3 && nc == 3) return launch_shape<3, 3>(which, flags, a, s);
  if (dim == 3 && nc == 4) return launch_shape<3, 4>(which, flags, a, s);
  return hipErrorInvalidValue;
}

}  // namespace gmc
