// gmc_patch.hpp -- patch-local assembly plan (host side).
//
// The elements are ordered along a Morton curve of their centroids and cut
// into patches of PE consecutive elements; one wave assembles one patch (4
// lanes per element).  Every BSR block and rhs node the patch touches gets a
// slot in LDS.  A block whose contributions all come from the patch is owned:
// the patch stores its finished value straight into the BSR values.  The
// patch's partial sums of the other (shared) blocks go to halo slots, sorted
// by destination block, and a small kernel sums each shared block's 2-4
// partials in patch order and stores it.  Every BSR value and rhs entry is
// stored exactly once.
//
// Sums stay in a fixed order without atomics: the kernel adds the element
// blocks of one local block index lb = (b, a) at a time, and elements of the
// patch that add into the same slot in the same step are separated into
// phases (their rank among those elements, in patch order), one after the
// other.
#pragma once

#include <cstdint>
#include <vector>

namespace gmc {

struct PatchPlan {
  int pe = 16;                       // elements per patch
  int n_patches = 0;
  std::vector<int> order;            // element ids in patch order
  std::vector<int> slot_ptr;         // [n_patches + 1] block slots of each patch
  std::vector<int> slot_dest;        // >= 0: owned BSR block; < 0: -(halo slot) - 1
  std::vector<int> node_ptr;         // [n_patches + 1] node slots of each patch
  std::vector<int> node_dest;        // >= 0: owned node; < 0: -(halo node slot) - 1
  std::vector<uint16_t> ecode;       // [E * nen^2] (patch order): local slot | phase << 12
  std::vector<uint16_t> ncode;       // [E * nen]   (patch order): local node slot | phase << 12
  int max_slots = 0, max_nodes = 0, max_phase = 0, max_nphase = 0;
  // halo: slots sorted by destination; shared block hb_dest[h] sums the slots
  // hb_ptr[h] .. hb_ptr[h+1]-1 (in patch order), nodes likewise
  std::vector<int> hb_ptr, hb_dest;
  std::vector<int> hn_ptr, hn_dest;
  long long n_owned_blocks = 0, n_owned_nodes = 0;
  // halo slots of the boundary integrals: boundary element i (of bd_elems)
  // adds its block (b, a) into halo block slot bd_hslot[i * (nen^2 + nen) +
  // b * nen + a] and its node b rhs into halo node slot
  // bd_hslot[i * (nen^2 + nen) + nen^2 + b]; these come after every patch's
  // slots in the halo sums
  std::vector<int> bd_hslot;
};

// Builds the plan; blkmap[e * nen^2 + b * nen + a] is the BSR block of element
// e's block (b, a); bd_elems (n_bd, ascending) the elements with boundary
// sides, whose blocks and nodes then stay shared (halo) and get halo slots
// of their own.  Returns false when a patch needs more than 4095 slots or
// more than 15 phases (the 12 + 4 bit codes).
bool build_patch_plan(int dim, int n_nodes, long long n_elems, const double* coords,
                      const int* conn, const int* blkmap, long long nnzb, int pe, PatchPlan& plan,
                      const int* bd_elems = nullptr, int n_bd = 0);

}  // namespace gmc
