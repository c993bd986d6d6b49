// gmc_patch.cpp -- construction of the patch-local assembly plan (host only).
#include "gmc_patch.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <utility>

namespace gmc {

namespace {

// spreads the low 21 bits of x to every third bit (3D) / the low 31 bits to
// every second bit (2D)
uint64_t spread3(uint64_t x) {
  x &= 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}
uint64_t spread2(uint64_t x) {
  x &= 0x7fffffff;
  x = (x | x << 16) & 0x0000ffff0000ffffull;
  x = (x | x << 8) & 0x00ff00ff00ff00ffull;
  x = (x | x << 4) & 0x0f0f0f0f0f0f0f0full;
  x = (x | x << 2) & 0x3333333333333333ull;
  x = (x | x << 1) & 0x5555555555555555ull;
  return x;
}

}  // namespace

bool build_patch_plan(int dim, int n_nodes, long long n_elems, const double* coords,
                      const int* conn, const int* rowptr, const int* blkmap, long long nnzb, int pe,
                      PatchPlan& P) {
  const int nen = dim + 1, nl = nen * nen;
  P = PatchPlan();
  P.pe = pe;
  // ---- Morton order of element centroids -------------------------------
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n_nodes; i++)
    for (int k = 0; k < dim; k++) {
      lo[k] = std::min(lo[k], coords[(long long)i * dim + k]);
      hi[k] = std::max(hi[k], coords[(long long)i * dim + k]);
    }
  const double levels = dim == 2 ? 2147483647.0 : 2097151.0;
  std::vector<std::pair<uint64_t, int>> key((size_t)n_elems);
  for (long long e = 0; e < n_elems; e++) {
    uint64_t code = 0;
    for (int k = 0; k < dim; k++) {
      double c = 0.0;
      for (int a = 0; a < nen; a++) c += coords[(long long)conn[e * nen + a] * dim + k];
      c /= nen;
      const double span = hi[k] - lo[k];
      const double u = span > 0.0 ? (c - lo[k]) / span : 0.0;
      const uint64_t qk = (uint64_t)std::llround(std::min(std::max(u, 0.0), 1.0) * levels);
      code |= (dim == 2 ? spread2(qk) : spread3(qk)) << k;
    }
    key[e] = {code, (int)e};
  }
  std::sort(key.begin(), key.end());
  P.order.resize(n_elems);
  for (long long i = 0; i < n_elems; i++) P.order[i] = key[i].second;
  key.clear();
  key.shrink_to_fit();

  // ---- row ownership: node n belongs to the patch holding most of the
  // elements that contain it (the earliest such patch on a tie); that patch
  // writes the whole BSR row n and rhs entry n, every other patch's
  // contribution goes to the halo.  Elements are visited in patch order, so
  // a node's candidate patches arrive in ascending runs.
  P.n_patches = (int)((n_elems + pe - 1) / pe);
  std::vector<int> owner(n_nodes, -1), best(n_nodes, 0), cur(n_nodes, -1), run(n_nodes, 0);
  for (long long i = 0; i < n_elems; i++) {
    const int p = (int)(i / pe);
    for (int b = 0; b < nen; b++) {
      const int n = conn[(long long)P.order[i] * nen + b];
      if (cur[n] != p) {
        cur[n] = p;
        run[n] = 0;
      }
      if (++run[n] > best[n]) {
        best[n] = run[n];
        owner[n] = p;
      }
    }
  }
  cur.clear();
  run.clear();
  best.clear();
  P.slot_ptr.assign(P.n_patches + 1, 0);
  P.node_ptr.assign(P.n_patches + 1, 0);
  P.ecode.resize((size_t)n_elems * nl);
  P.ncode.resize((size_t)n_elems * nen);
  // stamp-based local maps: block t -> local slot of the current patch
  std::vector<int> bstamp(nnzb, -1), bslot(nnzb, 0), nstamp(n_nodes, -1), nslot(n_nodes, 0);
  std::vector<int> own_nodes, own_blk, halo_blk, halo_nodes;
  // phase counters keyed by (lb, local slot), reset through the touched list
  std::vector<int> phase_cnt, touched;
  std::vector<std::pair<long long, int>> shared_b, shared_n;  // (dest, global slot index)
  for (int p = 0; p < P.n_patches; p++) {
    const long long i0 = (long long)p * pe, i1 = std::min<long long>(i0 + pe, n_elems);
    own_nodes.clear();
    halo_nodes.clear();
    own_blk.clear();
    halo_blk.clear();
    for (long long i = i0; i < i1; i++) {
      const long long e = P.order[i];
      for (int b = 0; b < nen; b++) {
        const int n = conn[e * nen + b];
        if (nstamp[n] == p) continue;
        nstamp[n] = p;
        (owner[n] == p ? own_nodes : halo_nodes).push_back(n);
      }
    }
    std::sort(own_nodes.begin(), own_nodes.end());
    // slots: every block of the owned rows in BSR order, then the blocks of
    // other rows this patch adds to
    for (int n : own_nodes)
      for (int t = rowptr[n]; t < rowptr[n + 1]; t++) {
        bstamp[t] = p;
        bslot[t] = (int)own_blk.size();
        own_blk.push_back(t);
      }
    for (long long i = i0; i < i1; i++) {
      const long long e = P.order[i];
      for (int lb = 0; lb < nl; lb++) {
        const int t = blkmap[e * nl + lb];
        if (bstamp[t] == p) continue;
        bstamp[t] = p;
        halo_blk.push_back(t);
      }
    }
    for (size_t h = 0; h < halo_blk.size(); h++) bslot[halo_blk[h]] = (int)(own_blk.size() + h);
    // node slots: owned nodes (ascending), then the others
    for (size_t k = 0; k < own_nodes.size(); k++) nslot[own_nodes[k]] = (int)k;
    for (size_t k = 0; k < halo_nodes.size(); k++) nslot[halo_nodes[k]] = (int)(own_nodes.size() + k);
    const int ns = (int)(own_blk.size() + halo_blk.size());
    const int nn = (int)(own_nodes.size() + halo_nodes.size());
    if (ns > 4095 || nn > 4095) return false;
    P.max_slots = std::max(P.max_slots, ns);
    P.max_nodes = std::max(P.max_nodes, nn);
    // phases: rank of the element among the patch elements adding into the
    // same slot at the same local index
    const size_t need = (size_t)std::max(nl * ns, nen * nn);
    if (phase_cnt.size() < need) phase_cnt.resize(need, 0);
    touched.clear();
    for (long long i = i0; i < i1; i++) {
      const long long e = P.order[i];
      for (int lb = 0; lb < nl; lb++) {
        const int ls = bslot[blkmap[e * nl + lb]];
        const int key2 = lb * ns + ls;
        const int ph = phase_cnt[key2]++;
        if (ph == 0) touched.push_back(key2);
        if (ph > 15) return false;
        P.max_phase = std::max(P.max_phase, ph);
        P.ecode[(size_t)i * nl + lb] = (uint16_t)(ls | (ph << 12));
      }
    }
    for (int k : touched) phase_cnt[k] = 0;
    touched.clear();
    for (long long i = i0; i < i1; i++) {
      const long long e = P.order[i];
      for (int b = 0; b < nen; b++) {
        const int ln = nslot[conn[e * nen + b]];
        const int key2 = b * nn + ln;
        const int ph = phase_cnt[key2]++;
        if (ph == 0) touched.push_back(key2);
        if (ph > 15) return false;
        P.max_nphase = std::max(P.max_nphase, ph);
        P.ncode[(size_t)i * nen + b] = (uint16_t)(ln | (ph << 12));
      }
    }
    for (int k : touched) phase_cnt[k] = 0;
    // destinations
    const int sbase = (int)P.slot_dest.size(), nbase = (int)P.node_dest.size();
    for (int t : own_blk) P.slot_dest.push_back(t);
    P.n_owned_blocks += (long long)own_blk.size();
    for (size_t h = 0; h < halo_blk.size(); h++) {
      P.slot_dest.push_back(-1);
      shared_b.push_back({(long long)halo_blk[h], sbase + (int)(own_blk.size() + h)});
    }
    for (int n : own_nodes) P.node_dest.push_back(n);
    P.n_owned_nodes += (long long)own_nodes.size();
    for (size_t k = 0; k < halo_nodes.size(); k++) {
      P.node_dest.push_back(-1);
      shared_n.push_back({(long long)halo_nodes[k], nbase + (int)(own_nodes.size() + k)});
    }
    P.slot_ptr[p + 1] = (int)P.slot_dest.size();
    P.node_ptr[p + 1] = (int)P.node_dest.size();
  }
  // halo slots in patch-major order (a patch stores its partial sums as one
  // contiguous run); per destination, the list of its halo slots in patch
  // order, which is the order the halo kernel adds them in
  auto build_halo = [](std::vector<std::pair<long long, int>>& sh, std::vector<int>& dest,
                       std::vector<int>& ptr, std::vector<int>& hdest, std::vector<int>& hidx) {
    // sh holds (destination, global slot index) in patch order: the halo slot
    // of entry k is k
    std::vector<std::pair<long long, int>> byd(sh.size());
    for (size_t k = 0; k < sh.size(); k++) {
      dest[sh[k].second] = -(int)k - 1;
      byd[k] = {sh[k].first, (int)k};
    }
    std::sort(byd.begin(), byd.end());
    ptr.assign(1, 0);
    hdest.clear();
    hidx.resize(byd.size());
    for (size_t h = 0; h < byd.size(); h++) {
      if (h > 0 && byd[h].first != byd[h - 1].first) ptr.push_back((int)h);
      if (h == 0 || byd[h].first != byd[h - 1].first) hdest.push_back((int)byd[h].first);
      hidx[h] = byd[h].second;
    }
    if (!byd.empty()) ptr.push_back((int)byd.size());
  };
  build_halo(shared_b, P.slot_dest, P.hb_ptr, P.hb_dest, P.hb_idx);
  build_halo(shared_n, P.node_dest, P.hn_ptr, P.hn_dest, P.hn_idx);
  return true;
}

}  // namespace gmc
