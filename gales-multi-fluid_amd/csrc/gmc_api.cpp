// gmc_api.cpp -- C-ABI of the fluid_mc assembly library (host side).
//
// Builds once what the reference builds before its time loop (the FE graph of
// linear_system.hpp:40-66, u_ref_inv of fluid_2d.hpp:123-135), plus what the
// GPU path needs: a BSR pattern, the per-element block map, a greedy element
// colouring for the atomic-free scatter, and boundary-side lists.  Every
// numerical step of the assembly, solve and update runs in HIP kernels on the
// context's stream; the host only orchestrates and reads scalars.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gales_mc.h"
#include "gmc_devparams.hpp"
#include "gmc_kernels.hpp"
#include "gmc_patch.hpp"
#include "gmc_physics.hpp"

using gmc::DevParams;

struct gmc_ctx {
  gmc_params params;
  DevParams dp;
  int dim = 0, nc = 0, ndof = 0, nen = 0;
  int n_nodes = 0, n_elems = 0, n_sides = 0;
  long long nnzb = 0;
  std::vector<int> h_rowptr, h_colidx;
  std::vector<long long> color_off, bd_color_off;
  int n_bd_elems = 0;
  // device
  double *coords = nullptr, *I = nullptr, *P = nullptr, *vals = nullptr, *rhs = nullptr,
         *x = nullptr;
  float* vals32 = nullptr;
  int *conn = nullptr, *blkmap = nullptr, *items = nullptr, *bd_items = nullptr,
      *side_start = nullptr, *side_local = nullptr, *side_flag = nullptr, *rowptr = nullptr,
      *colidx = nullptr, *diag_blk = nullptr, *d_status = nullptr, *one_item = nullptr;
  unsigned* dmask = nullptr;
  DevParams* d_params = nullptr;
  unsigned long long* d_dtmin = nullptr;   // bits of the adaptive time step's minimum
  // element-buffer scatter: contributions of every BSR block / node row
  void* ebm = nullptr;
  double* ebr = nullptr;
  int *cptr = nullptr, *slot = nullptr, *nptr = nullptr, *nidx = nullptr;
  bool use_ebuf = false;
  // the element buffer holds FP32 entries only (the FP64 one did not fit):
  // FP32-tier assemblies use it, FP64 ones the colour scatter
  bool ebuf_fp32_only = false;
  // element chunks of the element buffer (1: one buffer for the whole mesh).
  // Chunk k holds elements chunk_off[k] .. chunk_off[k+1]-1 (index order)
  // with its own slot numbering, per-block ranges at cptr + k (nnzb + 1) and
  // boundary elements bd_sorted[bd_chunk_off[k] ..]; its gather adds into
  // the values the earlier chunks' gathers wrote.
  int n_chunks = 1;
  // FP32-tier interior elements take their least-squares term on the matrix
  // cores (GMC_FP32_MFMA=1; patch-local or element-buffer scatter, 2D, 2
  // components)
  bool fp32_mfma = false;
  std::vector<long long> chunk_off, bd_chunk_off;
  int* bd_sorted = nullptr;   // the boundary elements, ascending (chunks, patch halo)
  int* bd_hslot = nullptr;    // patch scatter: halo slots of the boundary elements
  // patch-local assembly (gmc_patch.hpp)
  bool use_patch = false;
  int last_scatter = 0;
  int *patch_items = nullptr, *slot_ptr = nullptr, *slot_dest = nullptr, *node_ptr = nullptr,
      *node_dest = nullptr, *hb_ptr = nullptr, *hb_dest = nullptr, *hn_ptr = nullptr,
      *hn_dest = nullptr;
  uint16_t *ecode = nullptr, *ncode = nullptr;
  double *halo_m = nullptr, *halo_r = nullptr;
  gmc::PatchArgs pa{};
  long long n_hb = 0, n_hn = 0;
  // GMRES workspace
  int gm_k = 0;
  double *V = nullptr, *w = nullptr, *z = nullptr, *dinv = nullptr, *part = nullptr,
         *hbuf = nullptr, *ybuf = nullptr, *xs = nullptr, *r64 = nullptr, *binv = nullptr,
         *I0 = nullptr, *dvec = nullptr;
  // block ILU(0) preconditioner (rocSPARSE bsrilu0 / bsrsv): the factors
  // overwrite a copy of the BSR values; the analyses depend on the pattern
  // only and are kept
  rocsparse_handle sp = nullptr;
  rocsparse_mat_descr sp_A = nullptr, sp_L = nullptr, sp_U = nullptr;
  rocsparse_mat_info sp_info = nullptr;
  double* ilu_vals = nullptr;
  void* sp_buf = nullptr;
  size_t sp_buf_bytes = 0;
  bool ilu_ready = false;
  // adjoint workspace
  double *gP = nullptr, *gT = nullptr, *ones_e = nullptr, *vpart = nullptr, *vout = nullptr,
         *hv = nullptr, *hgs = nullptr, *hgp = nullptr;
  // element outputs (gmc_element)
  double *out_m = nullptr, *out_r = nullptr;
  // magma solubility tables: the context's host copy (params points into it)
  // and its device copy, [p | T | H2O | CO2] per component
  std::vector<double> mg_host[GMC_MAX_COMP];
  double* mg_dev[GMC_MAX_COMP] = {};
  // magma point tables (gmc_magma.hip) and the element of every side
  double *mt = nullptr, *mtb = nullptr;
  int* side_elem_csr = nullptr;
  hipStream_t stream = nullptr, own_stream = nullptr;
  // device address of this device's allocation-free scratch (norm_scratch)
  double* scratch = nullptr;
  hipEvent_t ev[6] = {};
  gmc_timing timing{};
  bool have_system = false, have_solution = false, have_rhs = false;
  size_t dev_bytes = 0;
  std::string err;
};

static std::string g_create_error;

// GMRES(k): k up to the reference's maximum subspace of 200
// (belos_linear_solver_factory.hpp, "Num Blocks"); the dot-product partials
// and the Hessenberg column hold up to kMaxBasis vectors
static constexpr int kMaxRestart = 200;
static constexpr int kMaxBasis = 256;

// ---------------------------------------------------------------------------
// Scratch of gmc_rhs_norm_noalloc: the block partial sums of the norm's
// reduction (at most kScratchParts, the cap of gmc::dot_blocks) and the final
// sum, in a file-scope device array: 1280 doubles (10 KiB) on every device
// the library runs on, held until the process ends.  Host code never uses the
// array's name as an address: gmc_create takes its device address once per
// device (hipGetSymbolAddress, with that device current) and contexts keep it.
// ---------------------------------------------------------------------------
static constexpr int kScratchParts = 1024;
static constexpr int kScratchLen = kScratchParts + 256;
static __device__ __align__(256) double norm_scratch[kScratchLen];
static constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_of_current_device(double** out, std::string& err) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    err = "no usable current HIP device for the scratch array";
    return GMC_ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(norm_scratch));
    if (e != hipSuccess || !ptr) {
      err = std::string("hipGetSymbolAddress of the scratch array failed: ") + hipGetErrorString(e);
      return GMC_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return GMC_OK;
}

// Scatter mode of an assembly with these flags: 2 patch-local, 1 element
// buffer, 0 colour scatter.
static int scatter_of(const gmc_ctx* c, int flags) {
  if (c->use_patch) return 2;
  if (c->use_ebuf && (!c->ebuf_fp32_only || (flags & GMC_ASM_FP32))) return 1;
  return 0;
}

#define GMC_FAIL(ctx, code, msg)   \
  do {                             \
    (ctx)->err = (msg);            \
    return (code);                 \
  } while (0)

#define HIPC(ctx, call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) {                                                               \
      (ctx)->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call;     \
      return GMC_ERR_HIP;                                                                 \
    }                                                                                     \
  } while (0)

template <typename T>
static hipError_t dalloc(gmc_ctx* c, T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e == hipSuccess) c->dev_bytes += n * sizeof(T);
  return e;
}

static bool shape_supported(int dim, int nc) {
  return (dim == 2 && (nc == 2 || nc == 3)) || (dim == 3 && (nc == 3 || nc == 4));
}

// ---------------------------------------------------------------------------
// mesh preprocessing
// ---------------------------------------------------------------------------
template <int DIM>
static long long first_bad_element(const gmc_mesh& m) {
  const int nen = DIM + 1;
  for (long long e = 0; e < m.n_elems; e++) {
    double xe[12];
    for (int a = 0; a < nen; a++)
      for (int k = 0; k < DIM; k++) xe[a * DIM + k] = m.coords[(long long)m.conn[e * nen + a] * DIM + k];
    gmc::Geo<DIM> g;
    if (!gmc::geometry<DIM>(xe, g)) return e;
  }
  return -1;
}

// FE graph (linear_system.hpp:46-63) as BSR: row i = node i, columns = nodes
// sharing an element with it, sorted.
static void build_pattern(const gmc_mesh& m, int nen, std::vector<int>& rowptr,
                          std::vector<int>& colidx) {
  const int nn = m.n_nodes;
  std::vector<int> cnt(nn + 1, 0);
  for (long long e = 0; e < m.n_elems; e++)
    for (int a = 0; a < nen; a++) cnt[m.conn[e * nen + a] + 1]++;
  for (int i = 0; i < nn; i++) cnt[i + 1] += cnt[i];
  std::vector<int> inc(cnt[nn]);
  std::vector<int> pos(cnt.begin(), cnt.end() - 1);
  for (long long e = 0; e < m.n_elems; e++)
    for (int a = 0; a < nen; a++) inc[pos[m.conn[e * nen + a]]++] = (int)e;
  rowptr.assign(nn + 1, 0);
  colidx.clear();
  colidx.reserve((size_t)nn * (nen == 3 ? 7 : 15));
  std::vector<int> buf;
  for (int i = 0; i < nn; i++) {
    buf.clear();
    for (int t = cnt[i]; t < cnt[i + 1]; t++)
      for (int a = 0; a < nen; a++) buf.push_back(m.conn[(long long)inc[t] * nen + a]);
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    colidx.insert(colidx.end(), buf.begin(), buf.end());
    rowptr[i + 1] = (int)colidx.size();
  }
}

static int find_col(const std::vector<int>& rowptr, const std::vector<int>& colidx, int i, int j) {
  auto b = colidx.begin() + rowptr[i], e = colidx.begin() + rowptr[i + 1];
  auto it = std::lower_bound(b, e, j);
  if (it == e || *it != j) return -1;
  return (int)(it - colidx.begin());
}

// Greedy colouring in element order: an element takes the smallest colour not
// used by any element sharing one of its nodes.  Returns the colour count or
// -1 if more than 128 colours would be needed.
// Greedy colouring: elements that share a node get different colours.  Each
// node keeps a bit set of the colours used around it, grown by 64-bit words
// when a colour past the current width is needed (a node of valence k needs
// at least k colours).  Returns the number of colours.
static int greedy_colour(const gmc_mesh& m, int nen, const std::vector<int>& elems,
                         std::vector<int>& colour) {
  int words = 1;
  std::vector<uint64_t> used((size_t)m.n_nodes, 0);
  colour.assign(elems.size(), 0);
  int ncol = 0;
  for (size_t t = 0; t < elems.size(); t++) {
    const long long e = elems[t];
    int c = -1;
    for (int w = 0; w < words && c < 0; w++) {
      uint64_t f = 0;
      for (int a = 0; a < nen; a++) f |= used[(size_t)m.conn[e * nen + a] * words + w];
      if (~f) c = 64 * w + __builtin_ctzll(~f);
    }
    if (c < 0) {
      std::vector<uint64_t> wider((size_t)m.n_nodes * (words + 1), 0);
      for (long long n = 0; n < m.n_nodes; n++)
        for (int w = 0; w < words; w++) wider[(size_t)n * (words + 1) + w] = used[(size_t)n * words + w];
      used.swap(wider);
      c = 64 * words;
      words++;
    }
    colour[t] = c;
    ncol = std::max(ncol, c + 1);
    for (int a = 0; a < nen; a++)
      used[(size_t)m.conn[e * nen + a] * words + c / 64] |= 1ull << (c % 64);
  }
  return ncol;
}

static void order_by_colour(const std::vector<int>& elems, const std::vector<int>& colour,
                            int ncol, std::vector<int>& items, std::vector<long long>& off) {
  off.assign(ncol + 1, 0);
  for (int c : colour) off[c + 1]++;
  for (int c = 0; c < ncol; c++) off[c + 1] += off[c];
  std::vector<long long> pos(off.begin(), off.end() - 1);
  items.resize(elems.size());
  for (size_t t = 0; t < elems.size(); t++) items[pos[colour[t]]++] = elems[t];
}

// ---------------------------------------------------------------------------
// lifecycle
// ---------------------------------------------------------------------------
// Copies each magma component's tables into the context (the caller's arrays
// need not outlive gmc_create) and points params at the copies.
static int keep_tables(gmc_ctx* c, std::string& err) {
  for (int k = 0; k < c->params.nc; k++) {
    gmc_component& cp = c->params.comp[k];
    if (cp.kind != GMC_COMP_MAGMA) continue;
    gmc_magma& g = cp.magma;
    if (g.n_p < 2 || !g.p_table || !g.wf_h2o_l || !g.wf_co2_l || g.n_T < 0 || g.n_T == 1 ||
        (g.n_T >= 2 && !g.T_table)) {
      err = "magma needs a pressure table of n_p >= 2 (and n_T = 0 or >= 2 temperatures)";
      return GMC_ERR_ARG;
    }
    const size_t np = (size_t)g.n_p, nT = (size_t)g.n_T, nw = np * (nT ? nT : 1);
    std::vector<double>& h = c->mg_host[k];
    h.assign(g.p_table, g.p_table + np);
    if (nT) h.insert(h.end(), g.T_table, g.T_table + nT);
    h.insert(h.end(), g.wf_h2o_l, g.wf_h2o_l + nw);
    h.insert(h.end(), g.wf_co2_l, g.wf_co2_l + nw);
    g.p_table = h.data();
    g.T_table = nT ? h.data() + np : nullptr;
    g.wf_h2o_l = h.data() + np + nT;
    g.wf_co2_l = h.data() + np + nT + nw;
  }
  return GMC_OK;
}

// Points the device parameter block's table pointers at the device copies.
static void device_tables(const gmc_ctx* c, DevParams& dp) {
  for (int k = 0; k < c->params.nc; k++) {
    if (c->params.comp[k].kind != GMC_COMP_MAGMA) continue;
    const gmc_magma& g = c->params.comp[k].magma;
    const size_t np = (size_t)g.n_p, nT = (size_t)g.n_T, nw = np * (nT ? nT : 1);
    double* d = c->mg_dev[k];
    dp.mg[k].p_tab = d;
    dp.mg[k].T_tab = nT ? d + np : nullptr;
    dp.mg[k].wh = d + np + nT;
    dp.mg[k].wc = d + np + nT + nw;
  }
}

static bool has_magma(const gmc_params& p) {
  for (int k = 0; k < p.nc; k++)
    if (p.comp[k].kind == GMC_COMP_MAGMA) return true;
  return false;
}

static void free_all(gmc_ctx* c) {
  void* ptrs[] = {c->coords, c->I,    c->P,        c->vals,       c->rhs,       c->x,
                  c->vals32, c->conn, c->blkmap,   c->items,      c->bd_items,  c->side_start,
                  c->side_local, c->side_flag, c->rowptr, c->colidx, c->diag_blk, c->d_status,
                  c->one_item, c->dmask, c->d_params, c->V, c->w, c->z, c->dinv, c->part,
                  c->hbuf, c->ybuf, c->xs, c->r64, c->binv, c->I0, c->dvec, c->out_m, c->out_r, c->ebm, c->ebr, c->cptr,
                  c->slot, c->bd_sorted, c->bd_hslot, c->nptr, c->nidx, c->gP, c->gT, c->ones_e, c->vpart, c->vout, c->hv,
                  c->hgs, c->hgp, c->patch_items, c->slot_ptr, c->slot_dest, c->node_ptr,
                  c->node_dest, c->hb_ptr, c->hb_dest, c->hn_ptr, c->hn_dest, c->ecode, c->ncode,
                  c->halo_m, c->halo_r, c->d_dtmin, c->mg_dev[0], c->mg_dev[1], c->mg_dev[2],
                  c->mg_dev[3], c->mt, c->mtb, c->side_elem_csr};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->ilu_vals) (void)hipFree(c->ilu_vals);
  if (c->sp_buf) (void)hipFree(c->sp_buf);
  if (c->sp_info) (void)rocsparse_destroy_mat_info(c->sp_info);
  if (c->sp_A) (void)rocsparse_destroy_mat_descr(c->sp_A);
  if (c->sp_L) (void)rocsparse_destroy_mat_descr(c->sp_L);
  if (c->sp_U) (void)rocsparse_destroy_mat_descr(c->sp_U);
  if (c->sp) (void)rocsparse_destroy_handle(c->sp);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
}

extern "C" int gmc_create(const gmc_mesh* mesh, const gmc_params* params, gmc_ctx** out) {
  if (!out) return GMC_ERR_ARG;
  *out = nullptr;
  if (!mesh || !params) {
    g_create_error = "null mesh or params";
    return GMC_ERR_ARG;
  }
  const gmc_mesh& m = *mesh;
  const gmc_params& p = *params;
  if (p.dim != m.dim || (p.dim != 2 && p.dim != 3)) {
    g_create_error = "dim must be 2 or 3 and equal in mesh and params";
    return GMC_ERR_ARG;
  }
  if (!shape_supported(p.dim, p.nc)) {
    g_create_error = "supported (dim, nc): (2,2), (2,3), (3,3), (3,4)";
    return GMC_ERR_UNSUPPORTED;
  }
  if (p.tau_variant != GMC_TAU_NON_DIAG_COMP_2019 && p.tau_variant != GMC_TAU_NON_DIAG_COMP_2001 &&
      p.tau_variant != GMC_TAU_DIAG_INCOMP_2007 && p.tau_variant != GMC_TAU_DIAG_2014) {
    g_create_error = "tau_variant must be 2001, 2007, 2014 or 2019";
    return GMC_ERR_ARG;
  }
  if (p.tau_variant == GMC_TAU_NON_DIAG_COMP_2001 && p.steady_state) {
    // fluid_2d.hpp:900-904: lambda = 1/sqrt(0) + h/(2|v|) with steady_state
    // (alpha is never assigned), an infinite tau
    g_create_error = "tau_non_diag_comp_2001 with steady_state gives an infinite tau";
    return GMC_ERR_UNSUPPORTED;
  }
  if (p.dc_2006 && !(p.dc_sharp == 1.0 || p.dc_sharp == 1.5 || p.dc_sharp == 2.0)) {
    g_create_error = "dc_sharp must be 1, 1.5 or 2";
    return GMC_ERR_ARG;
  }
  if (!(p.dt > 0.0)) {
    g_create_error = "dt must be positive";
    return GMC_ERR_ARG;
  }
  if (m.n_nodes <= 0 || m.n_elems <= 0 || m.n_sides < 0 || !m.coords || !m.conn ||
      !m.node_flag || (m.n_sides > 0 && (!m.side_elem || !m.side_local || !m.side_flag))) {
    g_create_error = "empty or incomplete mesh";
    return GMC_ERR_ARG;
  }
  const int nen = p.dim + 1;
  for (long long i = 0; i < (long long)m.n_elems * nen; i++)
    if (m.conn[i] < 0 || m.conn[i] >= m.n_nodes) {
      g_create_error = "connectivity out of range";
      return GMC_ERR_ARG;
    }
  for (int i = 0; i < m.n_nodes; i++)
    if (m.node_flag[i] < 0 || m.node_flag[i] >= GMC_MAX_FLAG) {
      g_create_error = "node flag out of range [0, 16)";
      return GMC_ERR_ARG;
    }
  for (int s = 0; s < m.n_sides; s++)
    if (m.side_elem[s] < 0 || m.side_elem[s] >= m.n_elems || m.side_local[s] < 0 ||
        m.side_local[s] > p.dim || m.side_flag[s] < 0 || m.side_flag[s] >= GMC_MAX_FLAG) {
      g_create_error = "boundary side out of range";
      return GMC_ERR_ARG;
    }
  const long long bad = (p.dim == 2) ? first_bad_element<2>(m) : first_bad_element<3>(m);
  if (bad >= 0) {
    g_create_error = "Jacobian is not positive at element " + std::to_string(bad) +
                     " (nodes must be anticlockwise)";
    return GMC_ERR_MESH;
  }

  gmc_ctx* c = new gmc_ctx();
  c->params = p;
  c->dim = p.dim;
  c->nc = p.nc;
  c->ndof = p.dim + 2 + p.nc - 1;
  c->nen = nen;
  c->n_nodes = m.n_nodes;
  c->n_elems = m.n_elems;
  c->n_sides = m.n_sides;
  std::string err;
  int st = keep_tables(c, err);
  if (!st) st = gmc::build_devparams(c->params, c->dp, err);
  if (st) {
    g_create_error = err;
    delete c;
    return st;
  }

  // pattern + block map
  build_pattern(m, nen, c->h_rowptr, c->h_colidx);
  c->nnzb = (long long)c->h_colidx.size();
  std::vector<int> blkmap((size_t)m.n_elems * nen * nen);
  for (long long e = 0; e < m.n_elems; e++)
    for (int b = 0; b < nen; b++)
      for (int a = 0; a < nen; a++)
        blkmap[(e * nen + b) * nen + a] =
            find_col(c->h_rowptr, c->h_colidx, m.conn[e * nen + b], m.conn[e * nen + a]);
  std::vector<int> diag(m.n_nodes);
  for (int i = 0; i < m.n_nodes; i++) diag[i] = find_col(c->h_rowptr, c->h_colidx, i, i);

  // element-buffer scatter: the contributions of every BSR block t occupy the
  // slots cptr[t] .. cptr[t+1]-1 in ascending element order; slot[e, b, a] is
  // where element e stores its block (b, a).  Per node, the (element, local
  // node) codes of the rhs gather.
  const int ndof_ = p.dim + 2 + p.nc - 1;
  const long long mslot = (long long)nen * nen * ndof_ * ndof_, rslot = (long long)nen * ndof_;
  const double eb_bytes = (double)m.n_elems * mslot * sizeof(double);
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  const char* env = getenv("GMC_SCATTER");
  const std::string mode = env ? env : "";
  // patch-local assembly by default when a patch's LDS fits; then the element
  // buffer when it fits in half the free memory; else the colour scatter
  // (default only in 2D: the 3D patch (16 tetrahedra, 51 KB of LDS, 6 phases)
  // measured slower than the element buffer on the 128^3 Kuhn box, 202 vs
  // 126 ms FP64 and 97 vs 67 ms in the FP32 tier, LDS limiting occupancy)
  gmc::PatchPlan plan;
  // the elements with boundary sides, ascending: the patch plan gives their
  // boundary integrals halo slots of their own (one launch, no colours)
  std::vector<int> bd_elems;
  {
    std::vector<char> has(m.n_elems, 0);
    for (int s = 0; s < m.n_sides; s++) has[m.side_elem[s]] = 1;
    for (int e = 0; e < m.n_elems; e++)
      if (has[e]) bd_elems.push_back(e);
  }
  if ((mode == "" && p.dim == 2) || mode == "patch") {
    // GMC_PATCH_ELEMS=16 or 64 takes that patch size only
    const char* pe_env = getenv("GMC_PATCH_ELEMS");
    const int pe_only = pe_env ? atoi(pe_env) : 0;
    for (const int pe : gmc::PATCH_SIZES) {
      if (pe_only && pe != pe_only) continue;
      if (!gmc::build_patch_plan(p.dim, m.n_nodes, m.n_elems, m.coords, m.conn, blkmap.data(),
                                 c->nnzb, pe, plan, bd_elems.data(), (int)bd_elems.size()))
        continue;
      gmc::PatchArgs t{};
      t.max_slots = plan.max_slots;
      t.max_nodes = plan.max_nodes;
      if (gmc::patch_lds_bytes(ndof_, t) <= gmc::PATCH_LDS_MAX) {
        c->use_patch = true;
        break;
      }
    }
  }
  // without the patch plan (3D, or forced), the element buffer when it fits
  // in half the free memory, in element chunks of that size when the whole
  // one does not (at most 16; GMC_EBUF_CHUNKS forces a count); both tiers
  // take the same scatter
  int chunks = (int)std::ceil(eb_bytes / std::max(0.5 * (double)free_b, 1.0));
  if (const char* ce = getenv("GMC_EBUF_CHUNKS")) chunks = std::max(chunks, atoi(ce));
  chunks = std::max(chunks, 1);
  c->use_ebuf = !c->use_patch && chunks <= 16 && (mode == "" || mode == "ebuf");
  if (c->use_ebuf) c->n_chunks = (int)std::min<long long>(chunks, std::max(m.n_elems, 1));
  // an FP32-entry buffer (half the bytes) for the precision tier alone
  // (GMC_SCATTER=ebuf32 forces it)
  if (!c->use_patch && !c->use_ebuf && 0.5 * eb_bytes < 0.5 * (double)free_b &&
      (mode == "" || mode == "ebuf32")) {
    c->use_ebuf = true;
    c->ebuf_fp32_only = true;
  }
  if (const char* me = getenv("GMC_FP32_MFMA"))
    c->fp32_mfma = atoi(me) != 0 && (c->use_ebuf || c->use_patch) && p.dim == 2 && p.nc == 2;
  std::vector<int> cptr, slot, nptr, nidx;
  long long max_chunk_slots = 0;
  if (c->use_ebuf) {
    // per chunk: the contributions of every block t in ascending element
    // order at slots cptr_k[t] .. cptr_k[t+1]-1 of the chunk's buffer
    const int nl = nen * nen;
    const long long nb1 = c->nnzb + 1;
    c->chunk_off.resize(c->n_chunks + 1);
    for (int k = 0; k <= c->n_chunks; k++) c->chunk_off[k] = (long long)m.n_elems * k / c->n_chunks;
    cptr.assign((size_t)c->n_chunks * nb1, 0);
    slot.resize(blkmap.size());
    for (int k = 0; k < c->n_chunks; k++) {
      int* cp = cptr.data() + (size_t)k * nb1;
      const long long i0 = c->chunk_off[k] * nl, i1 = c->chunk_off[k + 1] * nl;
      for (long long i = i0; i < i1; i++) cp[blkmap[i] + 1]++;
      for (long long t = 0; t < c->nnzb; t++) cp[t + 1] += cp[t];
      std::vector<int> pos(cp, cp + c->nnzb);
      for (long long i = i0; i < i1; i++) slot[i] = pos[blkmap[i]]++;
      max_chunk_slots = std::max<long long>(max_chunk_slots, cp[c->nnzb]);
    }
  }
  {
    // node codes: needed by the rhs gather and by the matrix-free product in
    // both scatter modes
    nptr.assign(m.n_nodes + 1, 0);
    for (long long k = 0; k < (long long)m.n_elems * nen; k++) nptr[m.conn[k] + 1]++;
    for (int i = 0; i < m.n_nodes; i++) nptr[i + 1] += nptr[i];
    nidx.resize((size_t)m.n_elems * nen);
    std::vector<int> npos(nptr.begin(), nptr.end() - 1);
    for (long long e = 0; e < m.n_elems; e++)
      for (int b = 0; b < nen; b++) nidx[npos[m.conn[e * nen + b]]++] = (int)(e * gmc::NODE_PER_ELEM_MAX + b);
  }

  // colouring of the interior elements, only for the colour scatter of FP64
  // assemblies; the patch and element-buffer scatters launch the elements in
  // index order (their sums do not depend on the launch order)
  std::vector<int> all(m.n_elems), colour, items;
  for (int e = 0; e < m.n_elems; e++) all[e] = e;
  if (!c->use_patch && (!c->use_ebuf || c->ebuf_fp32_only)) {
    const int ncol = greedy_colour(m, nen, all, colour);
    order_by_colour(all, colour, ncol, items, c->color_off);
  } else {
    items = all;
    c->color_off.clear();
  }

  // boundary sides grouped by element, boundary elements coloured separately
  std::vector<int> side_start(m.n_elems + 1, 0), sl(m.n_sides), sf(m.n_sides);
  for (int s = 0; s < m.n_sides; s++) side_start[m.side_elem[s] + 1]++;
  for (int e = 0; e < m.n_elems; e++) side_start[e + 1] += side_start[e];
  {
    std::vector<int> pos(side_start.begin(), side_start.end() - 1);
    for (int s = 0; s < m.n_sides; s++) {
      const int t = pos[m.side_elem[s]]++;
      sl[t] = m.side_local[s];
      sf[t] = m.side_flag[s];
    }
  }
  std::vector<int> bd, bdcol, bditems;
  for (int e = 0; e < m.n_elems; e++)
    if (side_start[e + 1] > side_start[e]) bd.push_back(e);
  c->n_bd_elems = (int)bd.size();
  const int nbcol = bd.empty() ? 0 : greedy_colour(m, nen, bd, bdcol);
  if (!bd.empty()) order_by_colour(bd, bdcol, nbcol, bditems, c->bd_color_off);
  else c->bd_color_off.assign(1, 0);
  if (c->n_chunks > 1) {
    c->bd_chunk_off.resize(c->n_chunks + 1);
    for (int k = 0; k <= c->n_chunks; k++)
      c->bd_chunk_off[k] = std::lower_bound(bd.begin(), bd.end(), c->chunk_off[k]) - bd.begin();
  }

  std::vector<unsigned> dmask(m.n_nodes);
  for (int i = 0; i < m.n_nodes; i++) dmask[i] = p.dirichlet_mask[m.node_flag[i]];

  // device buffers
  const long long nd = (long long)m.n_nodes * c->ndof;
  auto fail = [&](hipError_t e) {
    g_create_error = std::string("HIP error in gmc_create: ") + hipGetErrorString(e);
    free_all(c);
    delete c;
    return GMC_ERR_HIP;
  };
  hipError_t e;
#define A_(ptr, n)                                   \
  if ((e = dalloc(c, &(ptr), (size_t)(n))) != hipSuccess) return fail(e);
#define U_(dst, src, n)                                                                         \
  if ((e = hipMemcpy(dst, src, (size_t)(n) * sizeof(*(dst)), hipMemcpyHostToDevice)) != hipSuccess) \
    return fail(e);
  A_(c->coords, (long long)m.n_nodes * p.dim);
  A_(c->conn, (long long)m.n_elems * nen);
  A_(c->I, nd);
  A_(c->P, nd);
  A_(c->rhs, nd);
  A_(c->x, nd);
  A_(c->vals, c->nnzb * c->ndof * c->ndof);
  A_(c->blkmap, blkmap.size());
  A_(c->items, items.size());
  A_(c->bd_items, bditems.size());
  A_(c->side_start, side_start.size());
  A_(c->side_local, sl.size());
  A_(c->side_flag, sf.size());
  A_(c->rowptr, c->h_rowptr.size());
  A_(c->colidx, c->h_colidx.size());
  A_(c->diag_blk, diag.size());
  A_(c->dmask, dmask.size());
  A_(c->d_params, 1);
  A_(c->d_status, 1);
  A_(c->one_item, 1);
  A_(c->d_dtmin, 1);
  A_(c->out_m, (long long)c->nen * c->ndof * c->nen * c->ndof);
  A_(c->ebr, (long long)m.n_elems * rslot);
  A_(c->nptr, nptr.size());
  A_(c->nidx, nidx.size());
  if (c->use_ebuf) {
    {
      // FP64 entries (the FP32 tier uses the first half of each slot's
      // bytes), or FP32 entries only
      double* ebm64 = nullptr;
      A_(ebm64, c->ebuf_fp32_only ? ((long long)m.n_elems * mslot + 1) / 2
                                  : max_chunk_slots * (mslot / ((long long)nen * nen)));
      c->ebm = ebm64;
    }
    A_(c->cptr, cptr.size());
    A_(c->slot, slot.size());
    if (c->n_chunks > 1 && !bd.empty()) A_(c->bd_sorted, bd.size());
  }
  A_(c->out_r, (long long)c->nen * c->ndof);
  U_(c->coords, m.coords, (long long)m.n_nodes * p.dim);
  U_(c->conn, m.conn, (long long)m.n_elems * nen);
  U_(c->blkmap, blkmap.data(), blkmap.size());
  U_(c->items, items.data(), items.size());
  if (!bditems.empty()) U_(c->bd_items, bditems.data(), bditems.size());
  U_(c->side_start, side_start.data(), side_start.size());
  if (m.n_sides) {
    U_(c->side_local, sl.data(), sl.size());
    U_(c->side_flag, sf.data(), sf.size());
  }
  U_(c->rowptr, c->h_rowptr.data(), c->h_rowptr.size());
  U_(c->colidx, c->h_colidx.data(), c->h_colidx.size());
  U_(c->diag_blk, diag.data(), diag.size());
  U_(c->dmask, dmask.data(), dmask.size());
  for (int k = 0; k < p.nc; k++)
    if (!c->mg_host[k].empty()) {
      A_(c->mg_dev[k], c->mg_host[k].size());
      U_(c->mg_dev[k], c->mg_host[k].data(), c->mg_host[k].size());
    }
  device_tables(c, c->dp);
  U_(c->d_params, &c->dp, 1);
  if (has_magma(p)) {
    // point tables: interior [E][NQ][2 states][NC][vals], sides [S][NSQ][NC][vals]
    const long long per = (long long)p.nc * gmc::kMagmaVals;
    A_(c->mt, (long long)m.n_elems * nen * 2 * per);
    A_(c->mtb, (m.n_sides > 0 ? (long long)m.n_sides * p.dim : 1) * per);
    std::vector<int> se(m.n_sides > 0 ? m.n_sides : 1, 0);
    for (int e = 0; e < m.n_elems; e++)
      for (int t = side_start[e]; t < side_start[e + 1]; t++) se[t] = e;
    A_(c->side_elem_csr, se.size());
    U_(c->side_elem_csr, se.data(), se.size());
  }
  U_(c->nptr, nptr.data(), nptr.size());
  U_(c->nidx, nidx.data(), nidx.size());
  if (c->use_ebuf) {
    U_(c->cptr, cptr.data(), cptr.size());
    U_(c->slot, slot.data(), slot.size());
    if (c->bd_sorted) U_(c->bd_sorted, bd.data(), bd.size());
  }
  if (c->use_patch) {
    if (!bd.empty()) {
      A_(c->bd_sorted, bd.size());
      U_(c->bd_sorted, bd.data(), bd.size());
      A_(c->bd_hslot, plan.bd_hslot.size());
      U_(c->bd_hslot, plan.bd_hslot.data(), plan.bd_hslot.size());
    }
    A_(c->patch_items, plan.order.size());
    A_(c->slot_ptr, plan.slot_ptr.size());
    A_(c->slot_dest, plan.slot_dest.size());
    A_(c->node_ptr, plan.node_ptr.size());
    A_(c->node_dest, plan.node_dest.size());
    A_(c->hb_ptr, plan.hb_ptr.size());
    A_(c->hb_dest, plan.hb_dest.size());
    A_(c->hn_ptr, plan.hn_ptr.size());
    A_(c->hn_dest, plan.hn_dest.size());
    A_(c->ecode, plan.ecode.size());
    A_(c->ncode, plan.ncode.size());
    const long long nhs = plan.hb_ptr.empty() ? 0 : plan.hb_ptr.back();
    const long long nhn = plan.hn_ptr.empty() ? 0 : plan.hn_ptr.back();
    A_(c->halo_m, nhs * ndof_ * ndof_);
    A_(c->halo_r, nhn * ndof_);
    U_(c->patch_items, plan.order.data(), plan.order.size());
    U_(c->slot_ptr, plan.slot_ptr.data(), plan.slot_ptr.size());
    U_(c->slot_dest, plan.slot_dest.data(), plan.slot_dest.size());
    U_(c->node_ptr, plan.node_ptr.data(), plan.node_ptr.size());
    U_(c->node_dest, plan.node_dest.data(), plan.node_dest.size());
    if (!plan.hb_dest.empty()) {
      U_(c->hb_ptr, plan.hb_ptr.data(), plan.hb_ptr.size());
      U_(c->hb_dest, plan.hb_dest.data(), plan.hb_dest.size());
    }
    if (!plan.hn_dest.empty()) {
      U_(c->hn_ptr, plan.hn_ptr.data(), plan.hn_ptr.size());
      U_(c->hn_dest, plan.hn_dest.data(), plan.hn_dest.size());
    }
    U_(c->ecode, plan.ecode.data(), plan.ecode.size());
    U_(c->ncode, plan.ncode.data(), plan.ncode.size());
    c->n_hb = (long long)plan.hb_dest.size();
    c->n_hn = (long long)plan.hn_dest.size();
    gmc::PatchArgs& pa = c->pa;
    pa.ecode = c->ecode;
    pa.ncode = c->ncode;
    pa.slot_ptr = c->slot_ptr;
    pa.slot_dest = c->slot_dest;
    pa.node_ptr = c->node_ptr;
    pa.node_dest = c->node_dest;
    pa.halo_m = c->halo_m;
    pa.halo_r = c->halo_r;
    pa.n_patches = plan.n_patches;
    pa.pe = plan.pe;
    pa.max_slots = plan.max_slots;
    pa.max_nodes = plan.max_nodes;
    pa.max_phase = plan.max_phase;
    pa.max_nphase = plan.max_nphase;
  }
#undef A_
#undef U_
  if ((e = hipMemset(c->d_status, 0, sizeof(int))) != hipSuccess) return fail(e);
  if ((e = hipMemset(c->I, 0, nd * sizeof(double))) != hipSuccess) return fail(e);
  if ((e = hipMemset(c->P, 0, nd * sizeof(double))) != hipSuccess) return fail(e);
  if ((e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking)) != hipSuccess)
    return fail(e);
  c->stream = c->own_stream;
  {
    std::string serr;
    if (scratch_of_current_device(&c->scratch, serr) != GMC_OK) {
      g_create_error = serr;
      free_all(c);
      delete c;
      return GMC_ERR_HIP;
    }
  }
  for (auto& ev : c->ev)
    if ((e = hipEventCreate(&ev)) != hipSuccess) return fail(e);
  *out = c;
  return GMC_OK;
}

extern "C" int gmc_destroy(gmc_ctx* c) {
  if (!c) return GMC_OK;
  (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
  return GMC_OK;
}

extern "C" const char* gmc_last_error(const gmc_ctx* c) {
  return c ? c->err.c_str() : g_create_error.c_str();
}

extern "C" int gmc_get_info(const gmc_ctx* c, gmc_info* info) {
  if (!c || !info) return GMC_ERR_ARG;
  info->dim = c->dim;
  info->nc = c->nc;
  info->ndof = c->ndof;
  info->nen = c->nen;
  info->n_nodes = c->n_nodes;
  info->n_elems = c->n_elems;
  info->n_sides = c->n_sides;
  info->nnzb = c->nnzb;
  info->n_colors = c->color_off.empty() ? 0 : (int)c->color_off.size() - 1;
  info->n_bd_elems = c->n_bd_elems;
  info->n_bd_colors = (int)c->bd_color_off.size() - 1;
  info->device_bytes = c->dev_bytes;
  info->scatter = scatter_of(c, 0);
  info->ebuf_chunks = c->use_ebuf && !c->ebuf_fp32_only ? c->n_chunks : (c->use_ebuf ? 1 : 0);
  info->fp32_mfma = c->fp32_mfma ? 1 : 0;
  return GMC_OK;
}

extern "C" int gmc_set_stream(gmc_ctx* c, void* s) {
  if (!c) return GMC_ERR_ARG;
  c->stream = s ? (hipStream_t)s : c->own_stream;
  return GMC_OK;
}

extern "C" int gmc_set_dt(gmc_ctx* c, double dt) {
  if (!c) return GMC_ERR_ARG;
  if (!(dt > 0.0)) GMC_FAIL(c, GMC_ERR_ARG, "dt must be positive");
  c->params.dt = dt;
  c->dp.dt = dt;
  HIPC(c, hipMemcpyAsync(c->d_params, &c->dp, sizeof(DevParams), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

static int check_status(gmc_ctx* c);

// adaptive_time_criteria::f_mc + set_dt (adaptive_time_step.hpp:183-253,
// 380-394) on the I state, with the context's dt as the previous step
extern "C" int gmc_adaptive_dt(gmc_ctx* c, double cfl, double dt_min, double dt_max, double* dt) {
  if (!c || !dt) return GMC_ERR_ARG;
  if (!(cfl > 0.0) || !(dt_min > 0.0) || !(dt_max >= dt_min))
    GMC_FAIL(c, GMC_ERR_ARG, "need cfl > 0 and 0 < dt_min <= dt_max");
  const double dt_prev = c->params.dt;
  const unsigned long long inf_bits = 0x7FF0000000000000ull;
  HIPC(c, hipMemcpyAsync(c->d_dtmin, &inf_bits, sizeof(inf_bits), hipMemcpyHostToDevice, c->stream));
  gmc::ElementDtArgs a{};
  a.coords = c->coords;
  a.conn = c->conn;
  a.I = c->I;
  a.params = c->d_params;
  a.n_elems = c->n_elems;
  a.cfl = cfl;
  a.dt_prev = dt_prev;
  a.gmin_bits = c->d_dtmin;
  a.status = c->d_status;
  HIPC(c, gmc::launch_element_dt(c->dim, c->nc, a, c->stream));
  unsigned long long bits = 0;
  HIPC(c, hipMemcpyAsync(&bits, c->d_dtmin, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
  const int st = check_status(c);
  if (st) return st;
  double gmin;
  std::memcpy(&gmin, &bits, sizeof(gmin));
  // at most 1 % growth, then std::max(std::min(dt, dt_max), dt_min)
  double t = (gmin > 1.01 * dt_prev) ? 1.01 * dt_prev : gmin;
  t = (dt_max < t) ? dt_max : t;
  t = (t < dt_min) ? dt_min : t;
  *dt = t;
  return gmc_set_dt(c, t);
}

// pressure_profile_mc::Runge_Kutta (pressure_profile.hpp:487-501) with the
// library's mixture properties: depth steps of h from y_top, dp/ds = 9.81 rho
extern "C" int gmc_hydrostatic_profile(const gmc_params* p, double p_top, double y_top, double h,
                                       int n_steps, const double* T, const double* T_half,
                                       const double* Y, const double* Y_half, double* y_out,
                                       double* p_out) {
  if (!p || !T || !T_half || !Y || !Y_half || !y_out || !p_out || n_steps < 1 || !(h > 0.0))
    return GMC_ERR_ARG;
  DevParams dp;
  std::string err;
  const int bst = gmc::build_devparams(*p, dp, err);
  if (bst != GMC_OK) return bst;
  const int nc = p->nc;
  int pst = 0;   // property status bits over the whole profile
  auto rho_of = [&](double pr_, double T_, const double* Yd) {
    double Yf[GMC_MAX_COMP], s = 0.0;
    for (int c = 0; c < nc - 1; c++) {
      Yf[c] = Yd[c];
      s += Yd[c];
    }
    Yf[nc - 1] = 1.0 - s;
    double rho = 0.0;
    switch (nc) {
      case 2: { gmc::Material<double> M; gmc::load_material<double, 2>(dp, M); gmc::Props<double, 2> o;
                gmc::properties<double, 2>(dp, M, pr_, T_, Yf, o); rho = o.rho; pst |= o.st; break; }
      case 3: { gmc::Material<double> M; gmc::load_material<double, 3>(dp, M); gmc::Props<double, 3> o;
                gmc::properties<double, 3>(dp, M, pr_, T_, Yf, o); rho = o.rho; pst |= o.st; break; }
      default: { gmc::Material<double> M; gmc::load_material<double, 4>(dp, M); gmc::Props<double, 4> o;
                 gmc::properties<double, 4>(dp, M, pr_, T_, Yf, o); rho = o.rho; pst |= o.st; break; }
    }
    return rho;
  };
  p_out[0] = p_top;
  for (int i = 0; i <= n_steps; i++) y_out[i] = y_top - h * i;
  for (int i = 0; i < n_steps; i++) {
    const double* A = Y + (size_t)i * (nc - 1);
    const double* A1 = Y_half + (size_t)i * (nc - 1);
    const double* A2 = Y + (size_t)(i + 1) * (nc - 1);
    const double m1 = 9.81 * rho_of(p_out[i], T[i], A);
    const double m2 = 9.81 * rho_of(p_out[i] + 0.5 * h * m1, T_half[i], A1);
    const double m3 = 9.81 * rho_of(p_out[i] + 0.5 * h * m2, T_half[i], A1);
    const double m4 = 9.81 * rho_of(p_out[i] + h * m3, T[i + 1], A2);
    p_out[i + 1] = p_out[i] + h / 6. * (m1 + 2.0 * (m2 + m3) + m4);
  }
  if (pst & gmc::kStTable) return GMC_ERR_RANGE;
  if (pst) return GMC_ERR_NUMERIC;
  return GMC_OK;
}

extern "C" int gmc_set_material(gmc_ctx* c, const double theta[6]) {
  if (!c || !theta) return GMC_ERR_ARG;
  gmc_params p = c->params;
  if (p.comp[0].kind == GMC_COMP_MAGMA || p.comp[1].kind == GMC_COMP_MAGMA)
    GMC_FAIL(c, GMC_ERR_UNSUPPORTED, "theta parameterises custom and ideal-gas components only");
  for (int k = 0; k < 2; k++) {
    gmc_component& cp = p.comp[k];
    const double r = theta[3 * k], mu = theta[3 * k + 1], cpv = theta[3 * k + 2];
    if (cp.kind == GMC_COMP_CUSTOM) {
      cp.rho = r;
      cp.mu = mu;
      cp.cp = cpv;
    } else {
      // theta = (R, mu, cp) of the ideal gas: molar_mass = 8.314462/R, gamma = cp/(cp - R)
      if (!(cpv > r) || !(r > 0.0)) GMC_FAIL(c, GMC_ERR_ARG, "ideal gas needs cp > R > 0");
      cp.molar_mass = 8.314462 / r;
      cp.mu = mu;
      cp.gamma = cpv / (cpv - r);
    }
  }
  DevParams dp;
  std::string err;
  const int st = gmc::build_devparams(p, dp, err);
  if (st) GMC_FAIL(c, st, err);
  c->params = p;
  device_tables(c, dp);
  c->dp = dp;
  HIPC(c, hipMemcpyAsync(c->d_params, &c->dp, sizeof(DevParams), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

// ---------------------------------------------------------------------------
// state
// ---------------------------------------------------------------------------
extern "C" int gmc_set_state(gmc_ctx* c, const double* I, const double* P) {
  if (!c) return GMC_ERR_ARG;
  const size_t n = (size_t)c->n_nodes * c->ndof * sizeof(double);
  if (I) HIPC(c, hipMemcpyAsync(c->I, I, n, hipMemcpyHostToDevice, c->stream));
  if (P) HIPC(c, hipMemcpyAsync(c->P, P, n, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

extern "C" int gmc_set_state_device(gmc_ctx* c, const double* I, const double* P) {
  if (!c) return GMC_ERR_ARG;
  const size_t n = (size_t)c->n_nodes * c->ndof * sizeof(double);
  if (I) HIPC(c, hipMemcpyAsync(c->I, I, n, hipMemcpyDeviceToDevice, c->stream));
  if (P) HIPC(c, hipMemcpyAsync(c->P, P, n, hipMemcpyDeviceToDevice, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

extern "C" int gmc_get_state(const gmc_ctx* cc, double* I, double* P) {
  gmc_ctx* c = const_cast<gmc_ctx*>(cc);
  if (!c) return GMC_ERR_ARG;
  const size_t n = (size_t)c->n_nodes * c->ndof * sizeof(double);
  if (I) HIPC(c, hipMemcpyAsync(I, c->I, n, hipMemcpyDeviceToHost, c->stream));
  if (P) HIPC(c, hipMemcpyAsync(P, c->P, n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

// ---------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------
static gmc::AsmArgs base_args(gmc_ctx* c) {
  gmc::AsmArgs a{};
  a.coords = c->coords;
  a.conn = c->conn;
  a.I = c->I;
  a.P = c->P;
  a.dmask = c->dmask;
  a.blkmap = c->blkmap;
  a.vals = c->vals;
  a.rhs = c->rhs;
  a.side_start = c->side_start;
  a.side_local = c->side_local;
  a.side_flag = c->side_flag;
  a.params = c->d_params;
  a.status = c->d_status;
  a.ebm = c->ebm;
  a.slot = c->slot;
  a.ebr = c->ebr;
  a.mt = c->mt;
  a.mtb = c->mtb;
  return a;
}

// Fills the magma point tables from the current I and P (no-op without magma
// components); the element kernels launched after it on the stream read them.
static int magma_points(gmc_ctx* c) {
  if (!c->mt) return GMC_OK;
  gmc::MagmaPointArgs a{};
  a.coords = c->coords;
  a.conn = c->conn;
  a.I = c->I;
  a.P = c->P;
  a.params = c->d_params;
  a.n_elems = c->n_elems;
  a.side_elem = c->side_elem_csr;
  a.side_local = c->side_local;
  a.n_sides = c->n_sides;
  a.mt = c->mt;
  a.mtb = c->mtb;
  a.status = c->d_status;
  HIPC(c, gmc::launch_magma_points(c->dim, c->nc, a, c->stream));
  return GMC_OK;
}

static int check_status(gmc_ctx* c) {
  int st = 0;
  HIPC(c, hipMemcpyAsync(&st, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  if (st) {
    HIPC(c, hipMemsetAsync(c->d_status, 0, sizeof(int), c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    if (st & 1) GMC_FAIL(c, GMC_ERR_MESH, "Jacobian is not positive in an element");
    if (st & gmc::kStTable)
      GMC_FAIL(c, GMC_ERR_RANGE, "p or T outside a magma solubility table (p_index goes out of the vector scope)");
    if (st & gmc::kStModel)
      GMC_FAIL(c, GMC_ERR_NUMERIC, "magma viscosity model fails (zero denominator or no termination)");
    GMC_FAIL(c, GMC_ERR_NUMERIC, "LU factorization fails in inverse function (DUDY singular)");
  }
  return GMC_OK;
}

// Enqueues one assembly on the stream: with the element buffer, one launch
// over all interior elements, one over the boundary elements, and the gather;
// otherwise clear + interior colours + boundary colours.
static int enqueue_assembly(gmc_ctx* c, int flags, bool timed) {
  const int f = flags & (GMC_ASM_MATRIX | GMC_ASM_RHS | GMC_ASM_FP32);
  const long long nd = (long long)c->n_nodes * c->ndof;
  hipStream_t s = c->stream;
  gmc::AsmArgs a = base_args(c);
  if (const int ms = magma_points(c)) return ms;
  if (timed) HIPC(c, hipEventRecord(c->ev[0], s));
  const bool patch = scatter_of(c, f) == 2;
  c->last_scatter = scatter_of(c, f);
  if (patch) {
    // patch kernel (every BSR value and rhs entry written once: owned slots
    // directly, shared ones through the halo gather), the boundary integrals
    // into halo slots of their own (every block and node they touch is
    // shared), then the halo gather, which adds them after the patches'
    // partial sums
    if (timed) HIPC(c, hipEventRecord(c->ev[1], s));
    a.items = c->patch_items;
    a.n_items = c->n_elems;
    HIPC(c, gmc::launch_patch(c->dim, c->nc,
                              f | ((c->fp32_mfma && (f & GMC_ASM_FP32)) ? gmc::PATCH_MFMA : 0), a,
                              c->pa, s));
    if (timed) HIPC(c, hipEventRecord(c->ev[2], s));
    // the boundary elements, one launch, into their halo slots
    a.items = c->bd_sorted;
    a.n_items = c->n_bd_elems;
    a.halo_m = c->halo_m;
    a.halo_r = c->halo_r;
    a.hslot = c->bd_hslot;
    HIPC(c, gmc::launch_assemble(c->dim, c->nc, gmc::ASM_BOUNDARY | gmc::ASM_HALO, f, a, s));
    if (timed) HIPC(c, hipEventRecord(c->ev[3], s));
    if (f & GMC_ASM_MATRIX)
      HIPC(c, gmc::launch_halo_gather(c->n_hb, c->ndof * c->ndof, c->hb_ptr, c->hb_dest, c->halo_m,
                                      c->vals, s));
    if (f & GMC_ASM_RHS)
      HIPC(c, gmc::launch_halo_gather(c->n_hn, c->ndof, c->hn_ptr, c->hn_dest, c->halo_r, c->rhs, s));
    if (timed) HIPC(c, hipEventRecord(c->ev[4], s));
    return GMC_OK;
  }
  if (c->last_scatter == 1 && c->n_chunks > 1) {
    // chunk by chunk: interior elements, the chunk's boundary elements, the
    // chunk's matrix gather (the first writes the values, the others add in
    // chunk order); the rhs gather once at the end (the element residuals
    // are kept for the whole mesh).  With chunks the interior interval of
    // the timing also holds the earlier chunks' boundary passes and gathers.
    if (timed) HIPC(c, hipEventRecord(c->ev[1], s));
    const long long nb1 = c->nnzb + 1;
    const int mf = (c->fp32_mfma && (f & GMC_ASM_FP32)) ? gmc::ASM_MFMA : 0;
    for (int k = 0; k < c->n_chunks; k++) {
      a.items = c->items + c->chunk_off[k];
      a.n_items = c->chunk_off[k + 1] - c->chunk_off[k];
      HIPC(c, gmc::launch_assemble(c->dim, c->nc, gmc::ASM_EBUF | mf, f, a, s));
      if (timed && k == c->n_chunks - 1) HIPC(c, hipEventRecord(c->ev[2], s));
      a.items = c->bd_sorted + c->bd_chunk_off[k];
      a.n_items = c->bd_chunk_off[k + 1] - c->bd_chunk_off[k];
      HIPC(c, gmc::launch_assemble(c->dim, c->nc, gmc::ASM_EBUF | gmc::ASM_BOUNDARY, f, a, s));
      if (timed && k == c->n_chunks - 1) HIPC(c, hipEventRecord(c->ev[3], s));
      if (f & GMC_ASM_MATRIX)
        HIPC(c, gmc::launch_gather(c->ndof, c->nen, c->nnzb, c->n_nodes, c->cptr + k * nb1, c->nptr,
                                   c->nidx, c->ebm, c->ebr, c->vals, c->rhs,
                                   (f & ~GMC_ASM_RHS) | (k ? gmc::GATHER_ADD : 0), s));
    }
    if (f & GMC_ASM_RHS)
      HIPC(c, gmc::launch_gather(c->ndof, c->nen, c->nnzb, c->n_nodes, nullptr, c->nptr, c->nidx,
                                 nullptr, c->ebr, c->vals, c->rhs, f & ~GMC_ASM_MATRIX, s));
    if (timed) HIPC(c, hipEventRecord(c->ev[4], s));
    return GMC_OK;
  }
  if (c->last_scatter == 1) {
    if (timed) HIPC(c, hipEventRecord(c->ev[1], s));
    a.items = c->items;
    a.n_items = c->n_elems;
    const int mf = (c->fp32_mfma && (f & GMC_ASM_FP32)) ? gmc::ASM_MFMA : 0;
    HIPC(c, gmc::launch_assemble(c->dim, c->nc, gmc::ASM_EBUF | mf, f, a, s));
    if (timed) HIPC(c, hipEventRecord(c->ev[2], s));
    a.items = c->bd_items;
    a.n_items = c->n_bd_elems;
    HIPC(c, gmc::launch_assemble(c->dim, c->nc, gmc::ASM_EBUF | gmc::ASM_BOUNDARY, f, a, s));
    if (timed) HIPC(c, hipEventRecord(c->ev[3], s));
    HIPC(c, gmc::launch_gather(c->ndof, c->nen, c->nnzb, c->n_nodes, c->cptr, c->nptr,
                               c->nidx, c->ebm, c->ebr, c->vals, c->rhs, f, s));
    if (timed) HIPC(c, hipEventRecord(c->ev[4], s));
    return GMC_OK;
  }
  if (flags & GMC_ASM_MATRIX) HIPC(c, gmc::launch_fill(c->vals, c->nnzb * c->ndof * c->ndof, 0.0, s));
  if (flags & GMC_ASM_RHS) HIPC(c, gmc::launch_fill(c->rhs, nd, 0.0, s));
  if (timed) HIPC(c, hipEventRecord(c->ev[1], s));
  const int nint = (int)c->color_off.size() - 1;
  for (int k = 0; k < nint; k++) {
    a.items = c->items + c->color_off[k];
    a.n_items = c->color_off[k + 1] - c->color_off[k];
    HIPC(c, gmc::launch_assemble(c->dim, c->nc, 0, f, a, s));
  }
  if (timed) HIPC(c, hipEventRecord(c->ev[2], s));
  const int nbd = (int)c->bd_color_off.size() - 1;
  for (int k = 0; k < nbd; k++) {
    a.items = c->bd_items + c->bd_color_off[k];
    a.n_items = c->bd_color_off[k + 1] - c->bd_color_off[k];
    HIPC(c, gmc::launch_assemble(c->dim, c->nc, gmc::ASM_BOUNDARY, f, a, s));
  }
  if (timed) HIPC(c, hipEventRecord(c->ev[3], s));
  if (timed) HIPC(c, hipEventRecord(c->ev[4], s));
  return GMC_OK;
}

extern "C" int gmc_assemble(gmc_ctx* c, int flags, double* seconds) {
  if (!c) return GMC_ERR_ARG;
  if (!(flags & (GMC_ASM_MATRIX | GMC_ASM_RHS))) flags |= GMC_ASM_MATRIX | GMC_ASM_RHS;
  int st = enqueue_assembly(c, flags, true);
  if (st) return st;
  HIPC(c, hipEventSynchronize(c->ev[4]));
  float t01 = 0, t12 = 0, t23 = 0, t34 = 0, t04 = 0;
  HIPC(c, hipEventElapsedTime(&t01, c->ev[0], c->ev[1]));
  HIPC(c, hipEventElapsedTime(&t12, c->ev[1], c->ev[2]));
  HIPC(c, hipEventElapsedTime(&t23, c->ev[2], c->ev[3]));
  HIPC(c, hipEventElapsedTime(&t34, c->ev[3], c->ev[4]));
  HIPC(c, hipEventElapsedTime(&t04, c->ev[0], c->ev[4]));
  c->timing.assemble_s = t04 * 1e-3;
  c->timing.clear_s = t01 * 1e-3;
  c->timing.interior_s = t12 * 1e-3;
  const int sc = c->last_scatter;
  c->timing.boundary_s = t23 * 1e-3;
  c->timing.gather_s = t34 * 1e-3;
  c->timing.interior_launches =
      sc == 1 ? c->n_chunks : (sc ? 1 : (c->color_off.empty() ? 0 : (int)c->color_off.size() - 1));
  c->timing.boundary_launches = sc == 1 ? c->n_chunks : (sc == 2 ? 1 : (int)c->bd_color_off.size() - 1);
  if (seconds) *seconds = c->timing.assemble_s;
  st = check_status(c);
  if (st) return st;
  if (flags & GMC_ASM_MATRIX) c->have_system = true;
  if (flags & GMC_ASM_RHS) c->have_rhs = true;
  return GMC_OK;
}

extern "C" int gmc_get_timing(const gmc_ctx* c, gmc_timing* t) {
  if (!c || !t) return GMC_ERR_ARG;
  *t = c->timing;
  return GMC_OK;
}

static int device_norm(gmc_ctx* c, const double* v, double* out) {
  const long long n = (long long)c->n_nodes * c->ndof;
  const int nb = gmc::dot_blocks(n);
  if (!c->part) {
    HIPC(c, dalloc(c, &c->part, (size_t)nb * kMaxBasis));
    HIPC(c, dalloc(c, &c->hbuf, 2 * kMaxBasis));
  }
  HIPC(c, gmc::launch_multi_dot(v, 0, 1, v, n, c->part, c->stream));
  HIPC(c, gmc::launch_reduce_parts(c->part, 1, nb, c->hbuf, c->stream));
  double s = 0;
  HIPC(c, hipMemcpyAsync(&s, c->hbuf, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  *out = std::sqrt(s);
  return GMC_OK;
}

extern "C" int gmc_residual(gmc_ctx* c, double* norm) {
  if (!c) return GMC_ERR_ARG;
  int st = enqueue_assembly(c, GMC_ASM_RHS, false);
  if (st) return st;
  st = check_status(c);
  if (st) return st;
  c->have_rhs = true;
  if (norm) return device_norm(c, c->rhs, norm);
  return GMC_OK;
}

extern "C" int gmc_element(gmc_ctx* c, int e, int boundary, double* m, double* r) {
  if (!c) return GMC_ERR_ARG;
  if (e < 0 || e >= c->n_elems) GMC_FAIL(c, GMC_ERR_ARG, "element index out of range");
  const int ne = c->nen * c->ndof;
  HIPC(c, hipMemcpyAsync(c->one_item, &e, sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemsetAsync(c->out_m, 0, sizeof(double) * ne * ne, c->stream));
  HIPC(c, hipMemsetAsync(c->out_r, 0, sizeof(double) * ne, c->stream));
  gmc::AsmArgs a = base_args(c);
  if (const int ms = magma_points(c)) return ms;
  a.items = c->one_item;
  a.n_items = 1;
  a.out_m = c->out_m;
  a.out_r = c->out_r;
  const int prec = (c->params.precision == GMC_PREC_FP32) ? GMC_ASM_FP32 : 0;
  HIPC(c, gmc::launch_assemble(c->dim, c->nc, gmc::ASM_OUT | (boundary ? gmc::ASM_BOUNDARY : 0),
                               GMC_ASM_MATRIX | GMC_ASM_RHS | prec, a, c->stream));
  int st = check_status(c);
  if (st) return st;
  if (m) HIPC(c, hipMemcpyAsync(m, c->out_m, sizeof(double) * ne * ne, hipMemcpyDeviceToHost, c->stream));
  if (r) HIPC(c, hipMemcpyAsync(r, c->out_r, sizeof(double) * ne, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

extern "C" int gmc_get_pattern(const gmc_ctx* c, int* rowptr, int* colidx) {
  if (!c) return GMC_ERR_ARG;
  if (rowptr) std::memcpy(rowptr, c->h_rowptr.data(), sizeof(int) * c->h_rowptr.size());
  if (colidx) std::memcpy(colidx, c->h_colidx.data(), sizeof(int) * c->h_colidx.size());
  return GMC_OK;
}

extern "C" int gmc_get_system(const gmc_ctx* cc, double* vals, double* rhs) {
  gmc_ctx* c = const_cast<gmc_ctx*>(cc);
  if (!c) return GMC_ERR_ARG;
  if (vals)
    HIPC(c, hipMemcpyAsync(vals, c->vals, sizeof(double) * c->nnzb * c->ndof * c->ndof,
                           hipMemcpyDeviceToHost, c->stream));
  if (rhs)
    HIPC(c, hipMemcpyAsync(rhs, c->rhs, sizeof(double) * c->n_nodes * c->ndof,
                           hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

extern "C" int gmc_spmv(gmc_ctx* c, const double* x, double* y) {
  if (!c || !x || !y) return GMC_ERR_ARG;
  const long long n = (long long)c->n_nodes * c->ndof;
  if (!c->w) HIPC(c, dalloc(c, &c->w, n));
  if (!c->z) HIPC(c, dalloc(c, &c->z, n));
  HIPC(c, hipMemcpyAsync(c->z, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPC(c, gmc::launch_spmv(c->ndof, c->n_nodes, c->rowptr, c->colidx, c->vals, c->z, c->w, c->stream));
  HIPC(c, hipMemcpyAsync(y, c->w, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

// ---------------------------------------------------------------------------
// GMRES(k) + point Jacobi (right preconditioning, x0 = 0), CGS2 Arnoldi.
// Stands in for belos_linear_solver::execute (belos...:103-218).
// ---------------------------------------------------------------------------
static int ensure_gmres(gmc_ctx* c, int k) {
  const long long n = (long long)c->n_nodes * c->ndof;
  if (c->gm_k < k) {
    if (c->V) (void)hipFree(c->V);
    c->V = nullptr;
    HIPC(c, dalloc(c, &c->V, (size_t)n * (k + 1)));
    c->gm_k = k;
  }
  if (!c->w) HIPC(c, dalloc(c, &c->w, n));
  if (!c->z) HIPC(c, dalloc(c, &c->z, n));
  if (!c->dinv) HIPC(c, dalloc(c, &c->dinv, n));
  if (!c->part) {
    HIPC(c, dalloc(c, &c->part, (size_t)gmc::dot_blocks(n) * kMaxBasis));
    HIPC(c, dalloc(c, &c->hbuf, 2 * kMaxBasis));
  }
  if (!c->ybuf) HIPC(c, dalloc(c, &c->ybuf, kMaxBasis));
  if (!c->xs) HIPC(c, dalloc(c, &c->xs, n));
  return GMC_OK;
}

#define SPC(c, expr)                                                                        \
  do {                                                                                      \
    const rocsparse_status _s = (expr);                                                     \
    if (_s != rocsparse_status_success)                                                     \
      GMC_FAIL(c, GMC_ERR_HIP, std::string("rocSPARSE error ") + std::to_string((int)_s) + \
                                   " in " #expr);                                           \
  } while (0)

// Block ILU(0) of the assembled matrix on its own pattern (rocsparse_dbsrilu0
// on a copy of the BSR values, row-major ndof x ndof blocks).  The symbolic
// analyses (of the factorisation and of the two triangular solves) are made
// once per context; the numeric factorisation per call.  A zero pivot in a
// diagonal block gives GMC_ERR_NUMERIC.
static int ilu_factor(gmc_ctx* c) {
  const int mb = c->n_nodes, nnzb = (int)c->nnzb, bd = c->ndof;
  const rocsparse_direction dir = rocsparse_direction_row;
  const rocsparse_operation op = rocsparse_operation_none;
  if (!c->sp) {
    SPC(c, rocsparse_create_handle(&c->sp));
    SPC(c, rocsparse_create_mat_descr(&c->sp_A));
    SPC(c, rocsparse_create_mat_descr(&c->sp_L));
    SPC(c, rocsparse_set_mat_fill_mode(c->sp_L, rocsparse_fill_mode_lower));
    SPC(c, rocsparse_set_mat_diag_type(c->sp_L, rocsparse_diag_type_unit));
    SPC(c, rocsparse_create_mat_descr(&c->sp_U));
    SPC(c, rocsparse_set_mat_fill_mode(c->sp_U, rocsparse_fill_mode_upper));
    SPC(c, rocsparse_set_mat_diag_type(c->sp_U, rocsparse_diag_type_non_unit));
    SPC(c, rocsparse_create_mat_info(&c->sp_info));
    HIPC(c, dalloc(c, &c->ilu_vals, (size_t)c->nnzb * bd * bd));
  }
  SPC(c, rocsparse_set_stream(c->sp, c->stream));
  HIPC(c, hipMemcpyAsync(c->ilu_vals, c->vals, sizeof(double) * c->nnzb * bd * bd,
                         hipMemcpyDeviceToDevice, c->stream));
  if (!c->ilu_ready) {
    size_t b0 = 0, b1 = 0, b2 = 0;
    SPC(c, rocsparse_dbsrilu0_buffer_size(c->sp, dir, mb, nnzb, c->sp_A, c->ilu_vals, c->rowptr,
                                          c->colidx, bd, c->sp_info, &b0));
    SPC(c, rocsparse_dbsrsv_buffer_size(c->sp, dir, op, mb, nnzb, c->sp_L, c->ilu_vals, c->rowptr,
                                        c->colidx, bd, c->sp_info, &b1));
    SPC(c, rocsparse_dbsrsv_buffer_size(c->sp, dir, op, mb, nnzb, c->sp_U, c->ilu_vals, c->rowptr,
                                        c->colidx, bd, c->sp_info, &b2));
    const size_t bytes = std::max(b0, std::max(b1, b2));
    if (bytes > c->sp_buf_bytes) {
      if (c->sp_buf) (void)hipFree(c->sp_buf);
      c->sp_buf = nullptr;
      HIPC(c, hipMalloc(&c->sp_buf, bytes));
      c->sp_buf_bytes = bytes;
    }
    SPC(c, rocsparse_dbsrilu0_analysis(c->sp, dir, mb, nnzb, c->sp_A, c->ilu_vals, c->rowptr,
                                       c->colidx, bd, c->sp_info, rocsparse_analysis_policy_reuse,
                                       rocsparse_solve_policy_auto, c->sp_buf));
    SPC(c, rocsparse_dbsrsv_analysis(c->sp, dir, op, mb, nnzb, c->sp_L, c->ilu_vals, c->rowptr,
                                     c->colidx, bd, c->sp_info, rocsparse_analysis_policy_reuse,
                                     rocsparse_solve_policy_auto, c->sp_buf));
    SPC(c, rocsparse_dbsrsv_analysis(c->sp, dir, op, mb, nnzb, c->sp_U, c->ilu_vals, c->rowptr,
                                     c->colidx, bd, c->sp_info, rocsparse_analysis_policy_reuse,
                                     rocsparse_solve_policy_auto, c->sp_buf));
    c->ilu_ready = true;
  }
  SPC(c, rocsparse_dbsrilu0(c->sp, dir, mb, nnzb, c->sp_A, c->ilu_vals, c->rowptr, c->colidx, bd,
                            c->sp_info, rocsparse_solve_policy_auto, c->sp_buf));
  rocsparse_int pos = -1;
  const rocsparse_status zs = rocsparse_bsrilu0_zero_pivot(c->sp, c->sp_info, &pos);
  if (zs == rocsparse_status_zero_pivot)
    GMC_FAIL(c, GMC_ERR_NUMERIC, "block ILU(0): zero pivot in block row " + std::to_string(pos));
  SPC(c, zs);
  return GMC_OK;
}

// out = U^-1 L^-1 in (tmp: a work vector)
static int ilu_apply(gmc_ctx* c, const double* in, double* tmp, double* out) {
  const int mb = c->n_nodes, nnzb = (int)c->nnzb, bd = c->ndof;
  const double one = 1.0;
  SPC(c, rocsparse_dbsrsv_solve(c->sp, rocsparse_direction_row, rocsparse_operation_none, mb, nnzb,
                                &one, c->sp_L, c->ilu_vals, c->rowptr, c->colidx, bd, c->sp_info,
                                in, tmp, rocsparse_solve_policy_auto, c->sp_buf));
  SPC(c, rocsparse_dbsrsv_solve(c->sp, rocsparse_direction_row, rocsparse_operation_none, mb, nnzb,
                                &one, c->sp_U, c->ilu_vals, c->rowptr, c->colidx, bd, c->sp_info,
                                tmp, out, rocsparse_solve_policy_auto, c->sp_buf));
  return GMC_OK;
}

// Solves J x = b into xout (device), b on the device.
static int gmres_device(gmc_ctx* c, const gmc_solver& s, const double* b, double* xout,
                        gmc_solve_info* info) {
  const int k = s.restart;
  if (k < 1 || k > kMaxRestart) GMC_FAIL(c, GMC_ERR_ARG, "restart must be in [1, 200]");
  int st = ensure_gmres(c, k);
  if (st) return st;
  const long long n = (long long)c->n_nodes * c->ndof;
  const int nb = gmc::dot_blocks(n);
  hipStream_t str = c->stream;
  auto t0 = std::chrono::steady_clock::now();
  if (s.precond < 0 || s.precond > 3) GMC_FAIL(c, GMC_ERR_ARG, "precond must be 0, 1, 2 or 3");
  const bool blockj = s.precond == 2;
  const bool ilu = s.precond == GMC_PRECOND_ILU0;
  if (ilu) {
    st = ilu_factor(c);
    if (st) return st;
  } else if (blockj) {
    if (!c->binv) HIPC(c, dalloc(c, &c->binv, (size_t)c->n_nodes * c->ndof * c->ndof));
    HIPC(c, gmc::launch_block_jacobi_inv(c->ndof, c->n_nodes, c->diag_blk, c->vals, c->binv,
                                         c->d_status, str));
    st = check_status(c);
    if (st) return st;
  } else if (s.precond == 1) {
    HIPC(c, gmc::launch_jacobi_inv(c->ndof, c->n_nodes, c->diag_blk, c->vals, c->dinv, str));
  } else {
    HIPC(c, gmc::launch_fill(c->dinv, n, 1.0, str));
  }
  HIPC(c, gmc::launch_fill(xout, n, 0.0, str));
  double bnorm;
  st = device_norm(c, b, &bnorm);
  if (st) return st;
  std::vector<double> H((size_t)(k + 1) * k, 0.0), cs(k), sn(k), g(k + 1), y(k), hc(k + 2),
      hh(2 * (k + 1) + 1);
  int it = 0;
  double res = bnorm;
  if (bnorm > 0.0) {
    while (it < s.max_iters) {
      HIPC(c, gmc::launch_residual_vec(c->ndof, c->n_nodes, c->rowptr, c->colidx, c->vals, xout, b,
                                       c->w, str));
      double beta;
      st = device_norm(c, c->w, &beta);
      if (st) return st;
      res = beta;
      if (beta <= s.rel_tol * bnorm) break;
      HIPC(c, gmc::launch_scale_div_h(c->w, beta, c->V, n, str));
      std::fill(g.begin(), g.end(), 0.0);
      g[0] = beta;
      int j;
      bool conv = false;
      for (j = 0; j < k && it < s.max_iters; j++, it++) {
        double* vj = c->V + (long long)j * n;
        if (ilu) {
          st = ilu_apply(c, vj, c->xs, c->z);
          if (st) return st;
        } else if (blockj) {
          HIPC(c, gmc::launch_block_apply(c->ndof, c->n_nodes, c->binv, vj, c->z, false, str));
        } else {
          HIPC(c, gmc::launch_mul(c->dinv, vj, c->z, n, str));
        }
        HIPC(c, gmc::launch_spmv(c->ndof, c->n_nodes, c->rowptr, c->colidx, c->vals, c->z, c->w, str));
        for (int pass = 0; pass < 2; pass++) {
          double* hp = c->hbuf + pass * (k + 1);
          HIPC(c, gmc::launch_multi_dot(c->V, n, j + 1, c->w, n, c->part, str));
          HIPC(c, gmc::launch_reduce_parts(c->part, j + 1, nb, hp, str));
          HIPC(c, gmc::launch_multi_axpy(c->V, n, j + 1, hp, c->w, n, str));
        }
        double* hn2 = c->hbuf + 2 * (k + 1);
        HIPC(c, gmc::launch_multi_dot(c->w, 0, 1, c->w, n, c->part, str));
        HIPC(c, gmc::launch_reduce_parts(c->part, 1, nb, hn2, str));
        HIPC(c, hipMemcpyAsync(hh.data(), c->hbuf, sizeof(double) * (2 * (k + 1) + 1),
                               hipMemcpyDeviceToHost, str));
        HIPC(c, hipStreamSynchronize(str));
        for (int i = 0; i <= j; i++) hc[i] = hh[i] + hh[(k + 1) + i];
        const double hn = std::sqrt(hh[2 * (k + 1)]);
        hc[j + 1] = hn;
        // v_{j+1} = w / hn
        HIPC(c, gmc::launch_scale_div_h(c->w, hn, c->V + (long long)(j + 1) * n, n, str));
        for (int i = 0; i < j; i++) {
          const double t0v = cs[i] * hc[i] + sn[i] * hc[i + 1];
          hc[i + 1] = -sn[i] * hc[i] + cs[i] * hc[i + 1];
          hc[i] = t0v;
        }
        const double den = std::sqrt(hc[j] * hc[j] + hc[j + 1] * hc[j + 1]);
        cs[j] = (den != 0.0) ? hc[j] / den : 1.0;
        sn[j] = (den != 0.0) ? hc[j + 1] / den : 0.0;
        hc[j] = den;
        hc[j + 1] = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        for (int i = 0; i <= j; i++) H[(size_t)i * k + j] = hc[i];
        res = std::fabs(g[j + 1]);
        if (res <= s.rel_tol * bnorm) {
          j++;
          it++;
          conv = true;
          break;
        }
      }
      for (int i = j - 1; i >= 0; i--) {
        double sm = g[i];
        for (int cidx = i + 1; cidx < j; cidx++) sm -= H[(size_t)i * k + cidx] * y[cidx];
        y[i] = sm / H[(size_t)i * k + i];
      }
      HIPC(c, hipMemcpyAsync(c->ybuf, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, str));
      if (ilu) {
        // x += M^-1 (V y): V y into w, M^-1 into z (xs as the solves' middle)
        HIPC(c, gmc::launch_fill(c->w, n, 0.0, str));
        HIPC(c, gmc::launch_update_x(c->V, n, j, c->ybuf, nullptr, c->w, n, str));
        st = ilu_apply(c, c->w, c->xs, c->z);
        if (st) return st;
        HIPC(c, gmc::launch_axpy(1.0, c->z, xout, n, str));
      } else if (blockj) {
        HIPC(c, gmc::launch_fill(c->xs, n, 0.0, str));
        HIPC(c, gmc::launch_update_x(c->V, n, j, c->ybuf, nullptr, c->xs, n, str));
        HIPC(c, gmc::launch_block_apply(c->ndof, c->n_nodes, c->binv, c->xs, xout, true, str));
      } else {
        HIPC(c, gmc::launch_update_x(c->V, n, j, c->ybuf, c->dinv, xout, n, str));
      }
      if (conv) break;
    }
    HIPC(c, gmc::launch_residual_vec(c->ndof, c->n_nodes, c->rowptr, c->colidx, c->vals, xout, b,
                                     c->w, str));
    st = device_norm(c, c->w, &res);
    if (st) return st;
  }
  HIPC(c, hipStreamSynchronize(str));
  auto t1 = std::chrono::steady_clock::now();
  if (info) {
    info->iterations = it;
    info->abs_residual = res;
    info->rel_residual = (bnorm > 0.0) ? res / bnorm : 0.0;
    info->seconds = std::chrono::duration<double>(t1 - t0).count();
  }
  return GMC_OK;
}

extern "C" int gmc_solve(gmc_ctx* c, const gmc_solver* s, gmc_solve_info* info) {
  if (!c || !s) return GMC_ERR_ARG;
  if (!c->have_system) GMC_FAIL(c, GMC_ERR_STATE, "gmc_solve needs an assembled matrix");
  const int st = gmres_device(c, *s, c->rhs, c->x, info);
  if (st) return st;
  c->have_solution = true;
  return GMC_OK;
}

extern "C" int gmc_get_solution(const gmc_ctx* cc, double* delta) {
  gmc_ctx* c = const_cast<gmc_ctx*>(cc);
  if (!c || !delta) return GMC_ERR_ARG;
  HIPC(c, hipMemcpyAsync(delta, c->x, sizeof(double) * c->n_nodes * c->ndof, hipMemcpyDeviceToHost,
                         c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

extern "C" int gmc_newton_update(gmc_ctx* c) {
  if (!c) return GMC_ERR_ARG;
  if (!c->have_solution) GMC_FAIL(c, GMC_ERR_STATE, "gmc_newton_update needs a solved update");
  HIPC(c, gmc::launch_newton_update(c->I, c->x, c->n_nodes, c->ndof, c->dim + 2, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  c->have_solution = false;
  return GMC_OK;
}

extern "C" int gmc_corrector(gmc_ctx* c) {
  if (!c) return GMC_ERR_ARG;
  if (!c->have_solution) GMC_FAIL(c, GMC_ERR_STATE, "gmc_corrector needs a solved update");
  HIPC(c, gmc::launch_axpy(1.0, c->x, c->I, (long long)c->n_nodes * c->ndof, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  c->have_solution = false;
  return GMC_OK;
}

extern "C" int gmc_dof_trimmer(gmc_ctx* c, int dof_index, double lo, double hi) {
  if (!c) return GMC_ERR_ARG;
  if (dof_index < 0 || dof_index >= c->ndof) GMC_FAIL(c, GMC_ERR_ARG, "dof index out of range");
  HIPC(c, gmc::launch_trim(c->I, c->n_nodes, c->ndof, dof_index, lo, hi, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

extern "C" int gmc_rhs_norm(const gmc_ctx* cc, double* norm) {
  gmc_ctx* c = const_cast<gmc_ctx*>(cc);
  if (!c || !norm) return GMC_ERR_ARG;
  if (!c->have_rhs) GMC_FAIL(c, GMC_ERR_STATE, "gmc_rhs_norm needs an assembled rhs");
  return device_norm(c, c->rhs, norm);
}

// ||rhs||_2 as gmc_rhs_norm, with the reduction's partial sums in the
// file-scope scratch array instead of context memory allocated on first use:
// this call allocates nothing.  Every rhs length fits: dot_blocks caps the
// partial sums at kScratchParts.  Nothing in the scratch needs clearing (each
// partial is written before it is read).  Calls on one device share the
// array, so calls that overlap on one device race.
extern "C" int gmc_rhs_norm_noalloc(const gmc_ctx* cc, double* norm) {
  gmc_ctx* c = const_cast<gmc_ctx*>(cc);
  if (!c || !norm) return GMC_ERR_ARG;
  if (!c->have_rhs) GMC_FAIL(c, GMC_ERR_STATE, "gmc_rhs_norm_noalloc needs an assembled rhs");
  if (!c->scratch) GMC_FAIL(c, GMC_ERR_STATE, "the context has no scratch address");
  const long long n = (long long)c->n_nodes * c->ndof;
  const int nb = gmc::dot_blocks(n);
  if (nb > kScratchParts)
    GMC_FAIL(c, GMC_ERR_ARG, "the norm needs more partial sums than the fixed scratch holds (1024)");
  HIPC(c, gmc::launch_multi_dot(c->rhs, 0, 1, c->rhs, n, c->scratch, c->stream));
  HIPC(c, gmc::launch_reduce_parts(c->scratch, 1, nb, c->scratch + kScratchParts, c->stream));
  double s = 0.0;
  HIPC(c, hipMemcpyAsync(&s, c->scratch + kScratchParts, sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  *norm = std::sqrt(s);
  return GMC_OK;
}

extern "C" int gmc_predictor(gmc_ctx* c) {
  if (!c) return GMC_ERR_ARG;
  HIPC(c, hipMemcpyAsync(c->P, c->I, sizeof(double) * c->n_nodes * c->ndof, hipMemcpyDeviceToDevice,
                         c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

// ---------------------------------------------------------------------------
// matrix-free FP64 Jacobian product and the precision tier's refinement
// ---------------------------------------------------------------------------
// y = J64(I, P) x on the device: every element forms its FP64 matrix on the
// fly, multiplies its nodal slice of x (Dirichlet rows: avg_val x) and stores
// the product in its element slot; the node gather sums the slots in fixed
// element order.  The FP64 matrix is never stored.
static int jvp_device(gmc_ctx* c, const double* x, double* y) {
  gmc::AsmArgs a = base_args(c);
  if (const int ms = magma_points(c)) return ms;
  a.xvec = x;
  hipStream_t s = c->stream;
  a.items = c->items;
  a.n_items = c->n_elems;
  HIPC(c, gmc::launch_assemble(c->dim, c->nc, gmc::ASM_JVP, GMC_ASM_MATRIX, a, s));
  a.items = c->bd_items;
  a.n_items = c->n_bd_elems;
  HIPC(c, gmc::launch_assemble(c->dim, c->nc, gmc::ASM_JVP | gmc::ASM_BOUNDARY, GMC_ASM_MATRIX, a, s));
  HIPC(c, gmc::launch_gather(c->ndof, c->nen, c->nnzb, c->n_nodes, nullptr, c->nptr, c->nidx,
                             nullptr, c->ebr, nullptr, y, GMC_ASM_RHS, s));
  return check_status(c);
}

extern "C" int gmc_jvp(gmc_ctx* c, const double* x, double* y) {
  if (!c || !x || !y) return GMC_ERR_ARG;
  const long long n = (long long)c->n_nodes * c->ndof;
  if (!c->w) HIPC(c, dalloc(c, &c->w, n));
  if (!c->z) HIPC(c, dalloc(c, &c->z, n));
  HIPC(c, hipMemcpyAsync(c->z, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int st = jvp_device(c, c->z, c->w);
  if (st) return st;
  HIPC(c, hipMemcpyAsync(y, c->w, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

// Newton update in the precision tier.  rhs = -R(I) is assembled in FP64 and
// the working matrix J32 with FP32 element matrices (FP64 accumulation); the
// correction equation J64 x = rhs is then solved by mixed-precision iterative
// refinement:
//   r = rhs - J64 x   (FP64, matrix-free);  solve J32 d = r (GMRES);  x += d
// res_norms[k] = ||rhs - J64 x_k|| for k = 0..n_refine (x_0 = 0), so
// res_norms[k] / res_norms[0] is the relative Newton-equation residual of the
// update left in x.
extern "C" int gmc_refine(gmc_ctx* c, const gmc_solver* s, int n_refine, double* res_norms) {
  if (!c || !s) return GMC_ERR_ARG;
  if (n_refine < 1) GMC_FAIL(c, GMC_ERR_ARG, "n_refine must be >= 1");
  if (s->restart < 1 || s->restart > kMaxRestart) GMC_FAIL(c, GMC_ERR_ARG, "restart must be in [1, 200]");
  const long long n = (long long)c->n_nodes * c->ndof;
  hipStream_t str = c->stream;
  int st = ensure_gmres(c, s->restart);
  if (st) return st;
  if (!c->r64) HIPC(c, dalloc(c, &c->r64, n));
  if (!c->dvec) HIPC(c, dalloc(c, &c->dvec, n));
  if (!c->I0) HIPC(c, dalloc(c, &c->I0, n));
  st = gmc_assemble(c, GMC_ASM_MATRIX | GMC_ASM_RHS | GMC_ASM_FP32, nullptr);
  if (st) return st;
  double nrm;
  st = device_norm(c, c->rhs, &nrm);
  if (st) return st;
  if (res_norms) res_norms[0] = nrm;
  HIPC(c, gmc::launch_fill(c->x, n, 0.0, str));
  HIPC(c, hipMemcpyAsync(c->r64, c->rhs, n * sizeof(double), hipMemcpyDeviceToDevice, str));
  const char* venv = getenv("GMC_VERBOSE");
  const bool verbose = venv && venv[0] && venv[0] != '0';
  for (int k = 1; k <= n_refine; k++) {
    st = gmres_device(c, *s, c->r64, c->dvec, nullptr);
    if (st) return st;
    HIPC(c, gmc::launch_axpy(1.0, c->dvec, c->x, n, str));
    // r64 = rhs - J64 x  (I0 holds J64 x)
    st = jvp_device(c, c->x, c->I0);
    if (st) return st;
    HIPC(c, hipMemcpyAsync(c->r64, c->rhs, n * sizeof(double), hipMemcpyDeviceToDevice, str));
    HIPC(c, gmc::launch_axpy(-1.0, c->I0, c->r64, n, str));
    st = device_norm(c, c->r64, &nrm);
    if (st) return st;
    if (res_norms) res_norms[k] = nrm;
    // GMC_VERBOSE=1: one line per refinement pass on stderr
    if (verbose) {
      std::fprintf(stderr, "gmc_refine: pass %d of %d, ||rhs - J64 x|| = %.3e\n", k, n_refine, nrm);
      std::fflush(stderr);
    }
  }
  HIPC(c, hipStreamSynchronize(str));
  c->have_solution = true;
  return GMC_OK;
}

// ---------------------------------------------------------------------------
// adjoint
// ---------------------------------------------------------------------------
static int vjp_impl(gmc_ctx* c, const double* v, double* g_state, double* g_prev, double* g_params,
                    int mode) {
  if (mode != GMC_VJP_EXACT && mode != GMC_VJP_FROZEN) GMC_FAIL(c, GMC_ERR_ARG, "unknown vjp mode");
  if (mode == GMC_VJP_EXACT && has_magma(c->params))
    GMC_FAIL(c, GMC_ERR_UNSUPPORTED, "the exact vjp does not differentiate magma properties (use GMC_VJP_FROZEN)");
  hipStream_t s = c->stream;
  if (mode == GMC_VJP_FROZEN) {
    if (g_prev || g_params)
      GMC_FAIL(c, GMC_ERR_ARG, "the frozen mode gives g_state = J^T v only (g_prev, g_params must be NULL)");
    if (!g_state) return GMC_OK;
    gmc::AsmArgs a = base_args(c);
    if (const int ms = magma_points(c)) return ms;
    a.xvec = v;
    a.items = c->items;
    a.n_items = c->n_elems;
    HIPC(c, gmc::launch_assemble(c->dim, c->nc, gmc::ASM_VJPT, GMC_ASM_MATRIX, a, s));
    a.items = c->bd_items;
    a.n_items = c->n_bd_elems;
    HIPC(c, gmc::launch_assemble(c->dim, c->nc, gmc::ASM_VJPT | gmc::ASM_BOUNDARY, GMC_ASM_MATRIX, a, s));
    HIPC(c, gmc::launch_gather(c->ndof, c->nen, c->nnzb, c->n_nodes, nullptr, c->nptr, c->nidx,
                               nullptr, c->ebr, nullptr, g_state, GMC_ASM_RHS, s));
    return check_status(c);
  }
  const long long E = c->n_elems, ne = (long long)c->nen * c->ndof;
  if (!c->gP) HIPC(c, dalloc(c, &c->gP, (size_t)(E * ne)));
  if (!c->gT) HIPC(c, dalloc(c, &c->gT, (size_t)(6 * E)));
  if (!c->ones_e) {
    HIPC(c, dalloc(c, &c->ones_e, (size_t)E));
    HIPC(c, gmc::launch_fill(c->ones_e, E, 1.0, s));
  }
  if (!c->vpart) HIPC(c, dalloc(c, &c->vpart, (size_t)6 * 1024));
  gmc::VjpArgs a{};
  a.coords = c->coords;
  a.conn = c->conn;
  a.I = c->I;
  a.P = c->P;
  a.dmask = c->dmask;
  a.n_elems = E;
  a.side_start = c->side_start;
  a.side_local = c->side_local;
  a.side_flag = c->side_flag;
  a.params = c->d_params;
  a.status = c->d_status;
  a.v = v;
  a.gI = c->ebr;
  a.gP = c->gP;
  a.gT = c->gT;
  a.want_params = g_params != nullptr;
  a.items = c->items;
  a.n_items = E;
  HIPC(c, gmc::launch_vjp(c->dim, c->nc, false, a, s));
  if (g_prev) HIPC(c, gmc::launch_vjp_prev(c->dim, c->nc, a, s));
  a.items = c->bd_items;
  a.n_items = c->n_bd_elems;
  HIPC(c, gmc::launch_vjp(c->dim, c->nc, true, a, s));
  if (g_state)
    HIPC(c, gmc::launch_gather(c->ndof, c->nen, 0, c->n_nodes, nullptr, c->nptr, c->nidx, nullptr,
                               c->ebr, nullptr, g_state, GMC_ASM_RHS, s));
  if (g_prev)
    HIPC(c, gmc::launch_gather(c->ndof, c->nen, 0, c->n_nodes, nullptr, c->nptr, c->nidx, nullptr,
                               c->gP, nullptr, g_prev, GMC_ASM_RHS, s));
  if (g_params) {
    HIPC(c, gmc::launch_multi_dot(c->gT, E, 6, c->ones_e, E, c->vpart, s));
    HIPC(c, gmc::launch_reduce_parts(c->vpart, 6, gmc::dot_blocks(E), g_params, s));
  }
  return check_status(c);
}

extern "C" int gmc_vjp_device(gmc_ctx* c, const double* v, double* g_state, double* g_prev,
                              double* g_params, int mode) {
  if (!c || !v) return GMC_ERR_ARG;
  return vjp_impl(c, v, g_state, g_prev, g_params, mode);
}

extern "C" int gmc_vjp(gmc_ctx* c, const double* v, double* g_state, double* g_prev,
                       double* g_params, int mode) {
  if (!c || !v) return GMC_ERR_ARG;
  const long long n = (long long)c->n_nodes * c->ndof;
  if (!c->hv) HIPC(c, dalloc(c, &c->hv, n));
  if (!c->hgs) HIPC(c, dalloc(c, &c->hgs, n));
  if (!c->hgp) HIPC(c, dalloc(c, &c->hgp, n));
  if (!c->vout) HIPC(c, dalloc(c, &c->vout, 8));
  HIPC(c, hipMemcpyAsync(c->hv, v, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int st = vjp_impl(c, c->hv, g_state ? c->hgs : nullptr, g_prev ? c->hgp : nullptr,
                    g_params ? c->vout : nullptr, mode);
  if (st) return st;
  if (g_state) HIPC(c, hipMemcpyAsync(g_state, c->hgs, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (g_prev) HIPC(c, hipMemcpyAsync(g_prev, c->hgp, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (g_params) HIPC(c, hipMemcpyAsync(g_params, c->vout, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return GMC_OK;
}

// R = -rhs of (I, P) given in device memory, into device memory R.
extern "C" int gmc_residual_device(gmc_ctx* c, const double* I, const double* P, double* R) {
  if (!c || !R) return GMC_ERR_ARG;
  const long long n = (long long)c->n_nodes * c->ndof;
  if (I) HIPC(c, hipMemcpyAsync(c->I, I, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  if (P) HIPC(c, hipMemcpyAsync(c->P, P, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  int st = enqueue_assembly(c, GMC_ASM_RHS, false);
  if (st) return st;
  HIPC(c, gmc::launch_scale_div_h(c->rhs, -1.0, R, n, c->stream));
  return check_status(c);
}
