// gmc_kernels.hpp -- launch interfaces shared by the host layer and the HIP
// kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gales_mc.h"
#include "gmc_physics.hpp"

namespace gmc {

enum { ASM_BOUNDARY = 1, ASM_OUT = 2, ASM_EBUF = 4, ASM_JVP = 8, ASM_VJPT = 16, ASM_HALO = 32,
       ASM_MFMA = 64 };
enum { MODE_SCATTER = 0, MODE_DENSE = 1, MODE_EBUF = 2, MODE_JVP = 3, MODE_VJPT = 4, MODE_HALO = 5 };

// contribution codes of the rhs gather: element * 4 + local node
static constexpr int NODE_PER_ELEM_MAX = 4;

// element buffer: matrix part nen^2 blocks of ndof^2 per element, stored in
// block-sorted slots (element-matrix precision); rhs part nen*ndof per element
// (FP64, element-major)
template <int DIM, int NC>
struct EbLayout {
  static constexpr int N = DIM + 2 + NC - 1, NEN = DIM + 1;
  static constexpr int MS = NEN * NEN * N * N;
  static constexpr int RS = NEN * N;
};

// Arguments of the element kernels.  All pointers are device memory.
struct AsmArgs {
  const double* coords;     // [n_nodes*dim]
  const int* conn;          // [n_elems*nen]
  const double* I;          // [n_nodes*ndof] current iterate
  const double* P;          // [n_nodes*ndof] previous step
  const unsigned* dmask;    // [n_nodes] Dirichlet dof mask
  const int* items;         // element ids of this launch (one colour)
  long long n_items;
  const int* blkmap;        // [n_elems*nen*nen] BSR block of (row node b, col node a)
  double* vals;             // BSR values [nnzb*ndof*ndof]
  double* rhs;              // [n_nodes*ndof]
  const int* side_start;    // [n_elems+1] boundary sides by element
  const int* side_local;
  const int* side_flag;
  const DevParams* params;
  int* status;              // bit 0: J <= 0, bit 1: LU failure
  double* out_m;            // ASM_OUT: [n_items*ne*ne]
  double* out_r;            // ASM_OUT: [n_items*ne]
  void* ebm;                // ASM_EBUF: matrix part, block-sorted slots of ndof^2 (double or float)
  const int* slot;          // ASM_EBUF: [n_elems*nen*nen] slot of (element, row node b, col node a)
  double* ebr;              // ASM_EBUF / ASM_JVP: rhs part [n_elems*RS]
  const double* xvec;       // ASM_JVP / ASM_VJPT: x of y = J x or y = J^T x [n_nodes*ndof]
  // ASM_HALO (boundary elements of the patch scatter): item i writes its
  // blocks and node rhs into the halo slots hslot[i * (nen^2 + nen) ..]
  double* halo_m;           // [halo block slots * ndof^2]
  double* halo_r;           // [halo node slots * ndof]
  const int* hslot;
  // magma point tables (null without magma components): interior points
  // [n_elems][NQ][P state, I state][NC][kMagmaVals], boundary side points
  // [n_sides (side_start order)][NSQ][NC][kMagmaVals]
  const double* mt;
  const double* mtb;
};

// Magma point kernels (gmc_magma.hip): fill the point tables of AsmArgs from
// the I and P states (one thread per element point and per side point).
struct MagmaPointArgs {
  const double* coords;
  const int* conn;
  const double* I;
  const double* P;
  const DevParams* params;
  long long n_elems;
  const int* side_elem;     // [n_sides] element of each side (side_start order)
  const int* side_local;
  long long n_sides;
  double* mt;
  double* mtb;
  int* status;              // kStTable / kStModel bits
};
hipError_t launch_magma_points(int dim, int nc, const MagmaPointArgs& a, hipStream_t s);

// which: ASM_BOUNDARY selects execute_b, ASM_OUT writes dense element
// outputs, ASM_EBUF the element buffer, neither: colour scatter.  flags
// GMC_ASM_FP32 forms the element matrix in FP32 (residual stays FP64).
// exact vector-Jacobian products (gmc_adjoint.hip)
struct VjpArgs {
  const double* coords;
  const int* conn;
  const double* I;
  const double* P;
  const unsigned* dmask;
  const int* items;         // elements of this launch
  long long n_items;
  long long n_elems;        // stride of gT
  const int* side_start;
  const int* side_local;
  const int* side_flag;
  const DevParams* params;
  int* status;
  const double* v;          // adjoint vector [n_nodes*ndof]
  double* gI;               // [n_elems*nen*ndof] per-element state slots
  double* gP;               // [n_elems*nen*ndof] previous-step slots (interior only)
  double* gT;               // [6][n_elems] parameter slots
  int want_params;          // form the 6 parameter directions (gT)
};
hipError_t launch_vjp(int dim, int nc, bool boundary, const VjpArgs& a, hipStream_t s);
// g_P slots of every interior element (k_vjp_prev, FP64, one quad per element)
hipError_t launch_vjp_prev(int dim, int nc, const VjpArgs& a, hipStream_t s);

// CFL-adaptive time step (gmc_steps.hip): minimum over the elements of the
// f_mc criterion, as the bits of a positive double in *gmin_bits
struct ElementDtArgs {
  const double* coords;
  const int* conn;
  const double* I;
  const DevParams* params;
  long long n_elems;
  double cfl, dt_prev;
  unsigned long long* gmin_bits;   // set to the bits of +inf before the launch
  int* status;
};
hipError_t launch_element_dt(int dim, int nc, const ElementDtArgs& a, hipStream_t s);

hipError_t launch_assemble(int dim, int nc, int which, int flags, const AsmArgs& a,
                           hipStream_t s);

// Patch-local assembly (gmc_patch.hpp): one 64-lane workgroup (one wave) per
// patch of 16 elements (a.items in patch order, a.n_items elements in all).
struct PatchArgs {
  const uint16_t* ecode;    // [E * nen^2] local block slot | phase << 12
  const uint16_t* ncode;    // [E * nen]   local node slot | phase << 12
  const int* slot_ptr;      // [n_patches + 1]
  const int* slot_dest;     // >= 0 owned BSR block, < 0: -(halo slot) - 1
  const int* node_ptr;
  const int* node_dest;
  double* halo_m;           // [n_halo_block_slots * ndof^2]
  double* halo_r;           // [n_halo_node_slots * ndof]
  int n_patches;
  int pe;                   // elements per patch (4 lanes each: one workgroup of 4 pe lanes)
  int max_slots, max_nodes, max_phase, max_nphase;
};
// Elements per patch, tried in this order: the largest whose LDS fits in
// PATCH_LDS_MAX bytes per workgroup.  64 triangles (4 waves) share a
// quarter fewer blocks with other patches than 16 do, which halves the halo.
static constexpr int PATCH_SIZES[] = {64, 16};
static constexpr size_t PATCH_LDS_MAX = 64 * 1024;
// FP32-tier patch kernels with the least-squares term on the matrix cores
// stage its operands in the block-slot LDS (8,320 B per wave, before the
// slots are cleared), so that region holds at least that much
static constexpr size_t PATCH_MFMA_WAVE_BYTES = 8320;
__host__ __device__ inline size_t patch_acc_doubles(int nb, int max_slots, int threads) {
  const size_t blocks = (size_t)max_slots * nb, stage = (size_t)(threads / 64) * PATCH_MFMA_WAVE_BYTES / 8;
  return blocks > stage ? blocks : stage;
}
// LDS bytes of one patch workgroup (4 pe lanes)
inline size_t patch_lds_bytes(int ndof, const PatchArgs& pa, bool mfma = false) {
  const size_t acc = mfma ? patch_acc_doubles(ndof * ndof, pa.max_slots, 4 * pa.pe)
                          : (size_t)pa.max_slots * ndof * ndof;
  return (acc + (size_t)pa.max_nodes * ndof) * sizeof(double) +
         ((size_t)pa.max_slots + (size_t)pa.max_nodes) * sizeof(int);
}
hipError_t launch_patch(int dim, int nc, int flags, const AsmArgs& a, const PatchArgs& pa,
                        hipStream_t s);
// out[dest[h] * nb + j] = sum_{c in ptr[h] .. ptr[h+1]) halo[c * nb + j], in order
hipError_t launch_halo_gather(long long n_groups, int nb, const int* ptr, const int* dest,
                              const double* halo, double* out, hipStream_t s);

// sums the element buffer into FP64 BSR values / rhs (flags: GMC_ASM_MATRIX,
// GMC_ASM_RHS; GMC_ASM_FP32 when the matrix part holds floats; GATHER_ADD:
// add the matrix sums to the values instead of writing them, for the element
// chunks after the first)
static constexpr int GATHER_ADD = 1 << 16;
// launch_patch: the FP32 tier's least-squares term on the matrix cores
static constexpr int PATCH_MFMA = 1 << 17;
hipError_t launch_gather(int ndof, int nen, long long nnzb, long long n_nodes, const int* cptr,
                         const int* nptr, const int* nidx, const void* ebm,
                         const double* ebr, double* vals, double* rhs, int flags, hipStream_t s);

// ---- solver kernels (gmc_solver.hip) --------------------------------------
hipError_t launch_fill(double* x, long long n, double v, hipStream_t s);
hipError_t launch_spmv(int ndof, int n_nodes, const int* rowptr, const int* colidx,
                       const double* vals, const double* x, double* y, hipStream_t s);
// y = b - J x
hipError_t launch_residual_vec(int ndof, int n_nodes, const int* rowptr, const int* colidx,
                               const double* vals, const double* x, const double* b, double* y,
                               hipStream_t s);
hipError_t launch_jacobi_inv(int ndof, int n_nodes, const int* diag_blk, const double* vals,
                             double* dinv, hipStream_t s);
// block Jacobi: inverse diagonal blocks (status bit 2 on a singular block) and
// y (+)= Binv x
hipError_t launch_block_jacobi_inv(int ndof, int n_nodes, const int* diag_blk, const double* vals,
                                   double* binv, int* status, hipStream_t s);
hipError_t launch_block_apply(int ndof, int n_nodes, const double* binv, const double* x, double* y,
                              bool accumulate, hipStream_t s);
// partial dot products of w with the m vectors V[0..m) (stride ld) into
// part[m*nblk]; nblk = dot_blocks(n)
int dot_blocks(long long n);
hipError_t launch_multi_dot(const double* V, long long ld, int m, const double* w, long long n,
                            double* part, hipStream_t s);
// out[i] = sum_b part[i*nblk + b], fixed order (deterministic)
hipError_t launch_reduce_parts(const double* part, int m, int nblk, double* out, hipStream_t s);
// w -= sum_i h[i] V_i  (h in device memory)
hipError_t launch_multi_axpy(const double* V, long long ld, int m, const double* h, double* w,
                             long long n, hipStream_t s);
// y = alpha * x (alpha on the device: y = x / dnorm[0])
hipError_t launch_scale_div(const double* x, const double* dnorm, double* y, long long n,
                            hipStream_t s);
// y = x / d with a host scalar d (0 when d == 0)
hipError_t launch_scale_div_h(const double* x, double d, double* y, long long n, hipStream_t s);
// x += dinv .* (sum_i y[i] V_i)
hipError_t launch_update_x(const double* V, long long ld, int m, const double* y,
                           const double* dinv, double* x, long long n, hipStream_t s);
hipError_t launch_mul(const double* a, const double* b, double* y, long long n, hipStream_t s);
// I += d; clamp I[i*ndof + trim] to [0, 1]
hipError_t launch_newton_update(double* I, const double* d, long long n_nodes, int ndof,
                                int trim, hipStream_t s);
hipError_t launch_axpy(double alpha, const double* x, double* y, long long n, hipStream_t s);
// I[i*ndof + k] = max(lo, min(I[i*ndof + k], hi)) for every node i
hipError_t launch_trim(double* I, long long n_nodes, int ndof, int k, double lo, double hi,
                       hipStream_t s);
hipError_t launch_copy_f64_to_f32(const double* x, float* y, long long n, hipStream_t s);

}  // namespace gmc
