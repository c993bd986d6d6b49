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
                      const int* conn, const int* blkmap, long long nnzb, int pe, PatchPlan& P,
                      const int* bd_elems, int n_bd) {
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

  // ---- global contribution counts: a block (node) is owned by a patch when
  // every contribution to it comes from that patch
  P.n_patches = (int)((n_elems + pe - 1) / pe);
  std::vector<int> bcount(nnzb, 0), ncount(n_nodes, 0);
  for (long long k = 0; k < n_elems * nl; k++) bcount[blkmap[k]]++;
  for (long long k = 0; k < n_elems * nen; k++) ncount[conn[k]]++;
  // the boundary integrals are contributions from outside every patch
  for (int i = 0; i < n_bd; i++) {
    const long long e = bd_elems[i];
    for (int lb = 0; lb < nl; lb++) bcount[blkmap[e * nl + lb]]++;
    for (int b = 0; b < nen; b++) ncount[conn[e * nen + b]]++;
  }
  P.slot_ptr.assign(P.n_patches + 1, 0);
  P.node_ptr.assign(P.n_patches + 1, 0);
  P.ecode.resize((size_t)n_elems * nl);
  P.ncode.resize((size_t)n_elems * nen);
  // stamp-based local maps: block t -> local slot of the current patch
  std::vector<int> bstamp(nnzb, -1), bslot(nnzb, 0), nstamp(n_nodes, -1), nslot(n_nodes, 0);
  std::vector<int> blks, bcnt, nodes, ncnt, own_blk, sh_blk, own_nodes, sh_nodes;
  // phase counters keyed by (lb, local slot), reset through the touched list
  std::vector<int> phase_cnt, touched;
  std::vector<std::pair<long long, int>> shared_b, shared_n;  // (dest, global slot index)
  for (int p = 0; p < P.n_patches; p++) {
    const long long i0 = (long long)p * pe, i1 = std::min<long long>(i0 + pe, n_elems);
    blks.clear();
    bcnt.clear();
    nodes.clear();
    ncnt.clear();
    for (long long i = i0; i < i1; i++) {
      const long long e = P.order[i];
      for (int lb = 0; lb < nl; lb++) {
        const int t = blkmap[e * nl + lb];
        if (bstamp[t] != p) {
          bstamp[t] = p;
          bslot[t] = (int)blks.size();
          blks.push_back(t);
          bcnt.push_back(0);
        }
        bcnt[bslot[t]]++;
      }
      for (int b = 0; b < nen; b++) {
        const int n = conn[e * nen + b];
        if (nstamp[n] != p) {
          nstamp[n] = p;
          nslot[n] = (int)nodes.size();
          nodes.push_back(n);
          ncnt.push_back(0);
        }
        ncnt[nslot[n]]++;
      }
    }
    // slots: owned blocks in BSR order (their stores run along the rows),
    // then the shared ones; nodes likewise
    own_blk.clear();
    sh_blk.clear();
    for (size_t k = 0; k < blks.size(); k++)
      (bcnt[k] == bcount[blks[k]] ? own_blk : sh_blk).push_back(blks[k]);
    std::sort(own_blk.begin(), own_blk.end());
    for (size_t k = 0; k < own_blk.size(); k++) bslot[own_blk[k]] = (int)k;
    for (size_t k = 0; k < sh_blk.size(); k++) bslot[sh_blk[k]] = (int)(own_blk.size() + k);
    own_nodes.clear();
    sh_nodes.clear();
    for (size_t k = 0; k < nodes.size(); k++)
      (ncnt[k] == ncount[nodes[k]] ? own_nodes : sh_nodes).push_back(nodes[k]);
    std::sort(own_nodes.begin(), own_nodes.end());
    for (size_t k = 0; k < own_nodes.size(); k++) nslot[own_nodes[k]] = (int)k;
    for (size_t k = 0; k < sh_nodes.size(); k++) nslot[sh_nodes[k]] = (int)(own_nodes.size() + k);
    const int ns = (int)blks.size(), nn = (int)nodes.size();
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
    for (size_t h = 0; h < sh_blk.size(); h++) {
      P.slot_dest.push_back(-1);
      shared_b.push_back({(long long)sh_blk[h], sbase + (int)(own_blk.size() + h)});
    }
    for (int n : own_nodes) P.node_dest.push_back(n);
    P.n_owned_nodes += (long long)own_nodes.size();
    for (size_t k = 0; k < sh_nodes.size(); k++) {
      P.node_dest.push_back(-1);
      shared_n.push_back({(long long)sh_nodes[k], nbase + (int)(own_nodes.size() + k)});
    }
    P.slot_ptr[p + 1] = (int)P.slot_dest.size();
    P.node_ptr[p + 1] = (int)P.node_dest.size();
  }
  // the boundary elements' slots, after every patch's
  const int bd_b0 = (int)P.slot_dest.size(), bd_n0 = (int)P.node_dest.size();
  for (int i = 0; i < n_bd; i++) {
    const long long e = bd_elems[i];
    for (int lb = 0; lb < nl; lb++) {
      shared_b.push_back({(long long)blkmap[e * nl + lb], (int)P.slot_dest.size()});
      P.slot_dest.push_back(-1);
    }
    for (int b = 0; b < nen; b++) {
      shared_n.push_back({(long long)conn[e * nen + b], (int)P.node_dest.size()});
      P.node_dest.push_back(-1);
    }
  }
  // halo slots sorted by destination, then by patch (the global slot index
  // grows with the patch, the boundary elements' last): each shared block's
  // partial sums form one contiguous run, added in that order by the halo
  // kernel
  auto build_halo = [](std::vector<std::pair<long long, int>>& sh, std::vector<int>& dest,
                       std::vector<int>& ptr, std::vector<int>& hdest) {
    std::sort(sh.begin(), sh.end());
    ptr.assign(1, 0);
    hdest.clear();
    for (size_t h = 0; h < sh.size(); h++) {
      if (h > 0 && sh[h].first != sh[h - 1].first) ptr.push_back((int)h);
      if (h == 0 || sh[h].first != sh[h - 1].first) hdest.push_back((int)sh[h].first);
      dest[sh[h].second] = -(int)h - 1;
    }
    if (!sh.empty()) ptr.push_back((int)sh.size());
  };
  build_halo(shared_b, P.slot_dest, P.hb_ptr, P.hb_dest);
  build_halo(shared_n, P.node_dest, P.hn_ptr, P.hn_dest);
  P.bd_hslot.resize((size_t)n_bd * (nl + nen));
  for (int i = 0; i < n_bd; i++) {
    for (int lb = 0; lb < nl; lb++) P.bd_hslot[(size_t)i * (nl + nen) + lb] = -P.slot_dest[bd_b0 + i * nl + lb] - 1;
    for (int b = 0; b < nen; b++)
      P.bd_hslot[(size_t)i * (nl + nen) + nl + b] = -P.node_dest[bd_n0 + i * nen + b] - 1;
  }
  return true;
}

}  // namespace gmc
