// Test shim: compiles the product's patch planner (gmc_patch.cpp) for the CPU
// and emulates the patch kernel's summation schedule with integer-valued
// contributions (exact in any order), so tests can check the plan without a
// GPU: every element once, no two same-phase contributions to one slot, every
// BSR block and rhs node stored exactly once (by its owning patch or by the
// halo kernel), every boundary element's halo slot written once by that
// element alone, and the sums (boundary integrals included) equal a direct
// scatter.  Not part of the product.
#include <cstdint>
#include <vector>

#include "gmc_patch.cpp"

using namespace gmc;

static double contrib(long long e, int lb, int j) { return (double)((e * 7 + lb * 3 + j * 5) % 17) - 8.0; }

extern "C" int patch_check(int dim, int n_nodes, long long n_elems, const double* coords,
                           const int* conn, const int* blkmap, long long nnzb, int pe,
                           const int* bd_elems, int n_bd, long long* stats) {
  PatchPlan P;
  if (!build_patch_plan(dim, n_nodes, n_elems, coords, conn, blkmap, nnzb, pe, P, bd_elems, n_bd)) return 1;
  const int nen = dim + 1, nl = nen * nen, NB = 3;  // 3 values per block suffice
  // every element exactly once
  std::vector<int> seen(n_elems, 0);
  for (long long i = 0; i < n_elems; i++) seen[P.order[i]]++;
  for (long long e = 0; e < n_elems; e++)
    if (seen[e] != 1) return 2;
  std::vector<double> vals(nnzb * NB, 0.0), rhs((size_t)n_nodes * NB, 0.0);
  std::vector<int> vwrites(nnzb, 0), nwrites(n_nodes, 0);
  std::vector<double> halo_m((size_t)(P.hb_ptr.empty() ? 0 : P.hb_ptr.back()) * NB, 0.0);
  std::vector<double> halo_r((size_t)(P.hn_ptr.empty() ? 0 : P.hn_ptr.back()) * NB, 0.0);
  for (int p = 0; p < P.n_patches; p++) {
    const int s0 = P.slot_ptr[p], ns = P.slot_ptr[p + 1] - s0;
    const int n0 = P.node_ptr[p], nn = P.node_ptr[p + 1] - n0;
    std::vector<double> accm((size_t)ns * NB, 0.0), accr((size_t)nn * NB, 0.0);
    const long long i0 = (long long)p * pe, i1 = std::min<long long>(i0 + pe, n_elems);
    for (int b = 0; b < nen; b++)  // rhs: node phases
      for (int ph = 0; ph <= P.max_nphase; ph++) {
        std::vector<int> hit(nn, 0);
        for (long long i = i0; i < i1; i++) {
          const unsigned code = P.ncode[(size_t)i * nen + b];
          if ((int)(code >> 12) != ph) continue;
          const int ln = code & 0xFFF;
          if (ln >= nn) return 3;
          if (hit[ln]++) return 4;
          if (P.node_dest[n0 + ln] >= 0 && P.node_dest[n0 + ln] != conn[(long long)P.order[i] * nen + b]) return 5;
          for (int j = 0; j < NB; j++) accr[(size_t)ln * NB + j] += contrib(P.order[i], 100 + b, j);
        }
      }
    for (int lb = 0; lb < nl; lb++)
      for (int ph = 0; ph <= P.max_phase; ph++) {
        std::vector<int> hit(ns, 0);
        for (long long i = i0; i < i1; i++) {
          const unsigned code = P.ecode[(size_t)i * nl + lb];
          if ((int)(code >> 12) != ph) continue;
          const int ls = code & 0xFFF;
          if (ls >= ns) return 6;
          if (hit[ls]++) return 7;
          const long long e = P.order[i];
          if (P.slot_dest[s0 + ls] >= 0 && P.slot_dest[s0 + ls] != blkmap[e * nl + lb]) return 8;
          for (int j = 0; j < NB; j++) accm[(size_t)ls * NB + j] += contrib(e, lb, j);
        }
      }
    for (int s = 0; s < ns; s++) {
      const int d = P.slot_dest[s0 + s];
      for (int j = 0; j < NB; j++) {
        if (d >= 0) vals[(size_t)d * NB + j] = accm[(size_t)s * NB + j];
        else halo_m[(size_t)(-d - 1) * NB + j] = accm[(size_t)s * NB + j];
      }
      if (d >= 0) vwrites[d]++;
    }
    for (int s = 0; s < nn; s++) {
      const int d = P.node_dest[n0 + s];
      for (int j = 0; j < NB; j++) {
        if (d >= 0) rhs[(size_t)d * NB + j] = accr[(size_t)s * NB + j];
        else halo_r[(size_t)(-d - 1) * NB + j] = accr[(size_t)s * NB + j];
      }
      if (d >= 0) nwrites[d]++;
    }
  }
  // the boundary kernel: each boundary element into its own halo slots
  {
    std::vector<int> bhit(halo_m.size() / NB, 0), nhit(halo_r.size() / NB, 0);
    if ((long long)P.bd_hslot.size() != (long long)n_bd * (nl + nen)) return 13;
    for (int i = 0; i < n_bd; i++) {
      const long long e = bd_elems[i];
      for (int lb = 0; lb < nl; lb++) {
        const int h = P.bd_hslot[(size_t)i * (nl + nen) + lb];
        if (h < 0 || h >= (int)bhit.size() || bhit[h]++) return 14;
        for (int j = 0; j < NB; j++) halo_m[(size_t)h * NB + j] = contrib(e, 200 + lb, j);
      }
      for (int b = 0; b < nen; b++) {
        const int h = P.bd_hslot[(size_t)i * (nl + nen) + nl + b];
        if (h < 0 || h >= (int)nhit.size() || nhit[h]++) return 15;
        for (int j = 0; j < NB; j++) halo_r[(size_t)h * NB + j] = contrib(e, 300 + b, j);
      }
    }
  }
  for (size_t h = 0; h + 1 < P.hb_ptr.size(); h++) {
    const int t = P.hb_dest[h];
    vwrites[t]++;
    for (int j = 0; j < NB; j++) {
      double s = 0.0;
      for (int c = P.hb_ptr[h]; c < P.hb_ptr[h + 1]; c++) s += halo_m[(size_t)c * NB + j];
      vals[(size_t)t * NB + j] = s;
    }
  }
  for (size_t h = 0; h + 1 < P.hn_ptr.size(); h++) {
    const int n = P.hn_dest[h];
    nwrites[n]++;
    for (int j = 0; j < NB; j++) {
      double s = 0.0;
      for (int c = P.hn_ptr[h]; c < P.hn_ptr[h + 1]; c++) s += halo_r[(size_t)c * NB + j];
      rhs[(size_t)n * NB + j] = s;
    }
  }
  for (long long t = 0; t < nnzb; t++)
    if (vwrites[t] != 1) return 9;
  for (int n = 0; n < n_nodes; n++)
    if (nwrites[n] != 1) return 10;
  // direct scatter
  std::vector<double> vref(nnzb * NB, 0.0), rref((size_t)n_nodes * NB, 0.0);
  for (long long e = 0; e < n_elems; e++) {
    for (int lb = 0; lb < nl; lb++)
      for (int j = 0; j < NB; j++) vref[(size_t)blkmap[e * nl + lb] * NB + j] += contrib(e, lb, j);
    for (int b = 0; b < nen; b++)
      for (int j = 0; j < NB; j++) rref[(size_t)conn[e * nen + b] * NB + j] += contrib(e, 100 + b, j);
  }
  for (int i = 0; i < n_bd; i++) {
    const long long e = bd_elems[i];
    for (int lb = 0; lb < nl; lb++)
      for (int j = 0; j < NB; j++) vref[(size_t)blkmap[e * nl + lb] * NB + j] += contrib(e, 200 + lb, j);
    for (int b = 0; b < nen; b++)
      for (int j = 0; j < NB; j++) rref[(size_t)conn[e * nen + b] * NB + j] += contrib(e, 300 + b, j);
  }
  if (vref != vals) return 11;
  if (rref != rhs) return 12;
  stats[0] = P.n_patches;
  stats[1] = P.max_slots;
  stats[2] = P.max_nodes;
  stats[3] = P.max_phase;
  stats[4] = P.max_nphase;
  stats[5] = P.n_owned_blocks;
  stats[6] = P.hb_ptr.empty() ? 0 : P.hb_ptr.back();
  stats[7] = (long long)P.hb_dest.size();
  return 0;
}
