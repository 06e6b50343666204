// Symbolic analysis of the sparse path's Newton matrix (see sparse_plan.h).
#include "sparse_plan.h"

#include <algorithm>
#include <set>

namespace zrk {

namespace {

// Minimum-degree ordering of the species graph (symmetrised pattern, the
// temperature excluded: its full row and column are ordered last and cause
// no fill), on bitset adjacency: eliminating v joins its neighbours into a
// clique.  Ties go to the lowest index, so the ordering is deterministic.
std::vector<int> min_degree_order(int nsp, const std::vector<std::set<int>>& adj0) {
  const int W = (nsp + 63) / 64;
  std::vector<uint64_t> adj((size_t)nsp * W, 0);
  auto row = [&](int v) { return adj.data() + (size_t)v * W; };
  for (int v = 0; v < nsp; ++v)
    for (int a : adj0[v]) row(v)[a >> 6] |= 1ull << (a & 63);
  std::vector<int> deg(nsp);
  auto count = [&](int v) {
    int c = 0;
    for (int w = 0; w < W; ++w) c += __builtin_popcountll(row(v)[w]);
    return c;
  };
  for (int v = 0; v < nsp; ++v) deg[v] = count(v);
  std::vector<char> done(nsp, 0);
  std::vector<int> order;
  order.reserve(nsp);
  std::vector<int> nb;
  for (int it = 0; it < nsp; ++it) {
    int v = -1;
    for (int u = 0; u < nsp; ++u)
      if (!done[u] && (v < 0 || deg[u] < deg[v])) v = u;
    done[v] = 1;
    order.push_back(v);
    nb.clear();
    for (int w = 0; w < W; ++w)
      for (uint64_t b = row(v)[w]; b; b &= b - 1) nb.push_back(w * 64 + __builtin_ctzll(b));
    for (int a : nb) {
      uint64_t* ra = row(a);
      const uint64_t* rv = row(v);
      for (int w = 0; w < W; ++w) ra[w] |= rv[w];
      ra[a >> 6] &= ~(1ull << (a & 63));
      ra[v >> 6] &= ~(1ull << (v & 63));
      deg[a] = count(a);
    }
  }
  return order;
}

}  // namespace

bool build_sparse_plan(const Mechanism& M, SparsePlan* P, std::string* err) {
  const int nsp = M.nsp, N = nsp + 1;
  if (nsp < 1 || M.nstep < 1) {
    *err = "empty mechanism";
    return false;
  }
  // pattern of J in old indices: entry (row, col)
  std::vector<std::set<int>> rows(N);
  for (int i = 0; i < N; ++i) {
    rows[i].insert(i);
    rows[i].insert(nsp);    // temperature column
    rows[nsp].insert(i);    // temperature row
  }
  for (int s = 0; s < M.nstep; ++s)
    for (int k = M.reac_ptr[s]; k < M.reac_ptr[s + 1]; ++k) {
      const int c = M.reac_sp[k];
      for (int q = M.reac_ptr[s]; q < M.reac_ptr[s + 1]; ++q) rows[M.reac_sp[q]].insert(c);
      for (int q = M.prod_ptr[s]; q < M.prod_ptr[s + 1]; ++q) rows[M.prod_sp[q]].insert(c);
    }
  std::vector<std::set<int>> adj(nsp);
  for (int r = 0; r < nsp; ++r)
    for (int c : rows[r])
      if (c < nsp && c != r) {
        adj[r].insert(c);
        adj[c].insert(r);
      }
  std::vector<int> perm = min_degree_order(nsp, adj);
  perm.push_back(nsp);
  std::vector<int> iperm(N);
  for (int i = 0; i < N; ++i) iperm[perm[i]] = i;

  // symbolic LU without pivoting on the symmetrised pattern (A + A^T, so
  // that U's pattern is L's transposed): the pattern of row i of L is the
  // set of nodes reached from the entries of A's row i by walking up the
  // elimination tree until a node already marked for row i (the row-subtree
  // characterisation), O(|L|) in all
  std::vector<std::vector<int>> sym(N);
  for (int i = 0; i < N; ++i)
    for (int c : rows[perm[i]]) {
      const int j = iperm[c];
      if (j < i) sym[i].push_back(j);
      if (j > i) sym[j].push_back(i);
    }
  std::vector<int> parent(N, -1), anc(N, -1);
  for (int i = 0; i < N; ++i)
    for (int k : sym[i])
      for (int r = k; r != -1 && r < i;) {  // elimination tree with path compression
        const int next = anc[r];
        anc[r] = i;
        if (next == -1) {
          parent[r] = i;
          break;
        }
        r = next;
      }
  std::vector<int> mark(N, -1);
  std::vector<std::vector<int>> lrow(N), ucol(N);  // L(i, :) and, transposed, U(:, j)
  for (int i = 0; i < N; ++i) {
    mark[i] = i;
    for (int k : sym[i])
      for (int r = k; r != -1 && mark[r] != i; r = parent[r]) {
        mark[r] = i;
        lrow[i].push_back(r);
      }
    std::sort(lrow[i].begin(), lrow[i].end());
    for (int k : lrow[i]) ucol[k].push_back(i);  // U(k, i) by symmetry
  }
  std::vector<std::vector<int>> lu(N);
  for (int i = 0; i < N; ++i) {
    lu[i] = lrow[i];
    lu[i].push_back(i);
    lu[i].insert(lu[i].end(), ucol[i].begin(), ucol[i].end());  // ascending already
  }
  P->N = N;
  P->perm = perm;
  P->rp.assign(N + 1, 0);
  for (int i = 0; i < N; ++i) P->rp[i + 1] = P->rp[i] + (int)lu[i].size();
  P->nnz = P->rp[N];
  P->ci.resize(P->nnz);
  P->dpos.resize(N);
  P->max_row = 0;
  P->factor_flops = 0;
  for (int i = 0; i < N; ++i) {
    std::copy(lu[i].begin(), lu[i].end(), P->ci.begin() + P->rp[i]);
    P->dpos[i] = P->rp[i] + (int)(std::lower_bound(lu[i].begin(), lu[i].end(), i) - lu[i].begin());
    P->max_row = std::max(P->max_row, (int)lu[i].size());
  }
  for (int i = 0; i < N; ++i)
    for (int p = P->rp[i]; p < P->dpos[i]; ++p) {
      const int k = P->ci[p];
      P->factor_flops += P->rp[k + 1] - P->dpos[k] - 1;
    }
  // level sets: forward (row i after every row of its lower part), backward
  // (row i after every row of its upper part)
  std::vector<int> lev(N, 0);
  int nl = 0;
  for (int i = 0; i < N; ++i) {
    int l = 0;
    for (int p = P->rp[i]; p < P->dpos[i]; ++p) l = std::max(l, lev[P->ci[p]] + 1);
    lev[i] = l;
    nl = std::max(nl, l + 1);
  }
  auto bucket = [&](int nlev, std::vector<int>& ptr, std::vector<int>& out) {
    ptr.assign(nlev + 1, 0);
    for (int i = 0; i < N; ++i) ptr[lev[i] + 1]++;
    for (int l = 0; l < nlev; ++l) ptr[l + 1] += ptr[l];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    out.resize(N);
    for (int i = 0; i < N; ++i) out[fill[lev[i]]++] = i;
  };
  P->nlev_f = nl;
  bucket(nl, P->lf_ptr, P->lf_rows);
  nl = 0;
  for (int i = N - 1; i >= 0; --i) {
    int l = 0;
    for (int p = P->dpos[i] + 1; p < P->rp[i + 1]; ++p) l = std::max(l, lev[P->ci[p]] + 1);
    lev[i] = l;
    nl = std::max(nl, l + 1);
  }
  P->nlev_b = nl;
  bucket(nl, P->lb_ptr, P->lb_rows);

  // the trailing block [split, N) x [split, N): the largest one (at least 64
  // rows) whose share of LU entries is at least 90%
  {
    std::vector<long long> colcnt(N, 0);  // L entries (r > c) per column
    for (int i = 0; i < N; ++i)
      for (int p = P->rp[i]; p < P->dpos[i]; ++p) colcnt[P->ci[p]]++;
    long long cnt = 0;
    P->split = N;
    for (int i = N - 1; i >= 0; --i) {
      cnt += P->rp[i + 1] - P->dpos[i];  // U part of row i with the diagonal
      cnt += colcnt[i];                  // L part of column i (rows > i, all in the block)
      const long long nd = N - i;
      if (nd >= 64 && cnt * 10 >= nd * nd * 9) P->split = i;
    }
  }
  {
    const int s = P->split;
    std::vector<int> lv(s, 0);
    int nl2 = 0;
    for (int i = 0; i < s; ++i) {
      int l = 0;
      for (int p = P->rp[i]; p < P->dpos[i]; ++p) l = std::max(l, lv[P->ci[p]] + 1);
      lv[i] = l;
      nl2 = std::max(nl2, l + 1);
    }
    P->nlev_l = nl2;
    P->ll_ptr.assign(nl2 + 1, 0);
    for (int i = 0; i < s; ++i) P->ll_ptr[lv[i] + 1]++;
    for (int l = 0; l < nl2; ++l) P->ll_ptr[l + 1] += P->ll_ptr[l];
    std::vector<int> fill(P->ll_ptr.begin(), P->ll_ptr.end() - 1);
    P->ll_rows.assign(std::max(s, 1), 0);
    for (int i = 0; i < s; ++i) P->ll_rows[fill[lv[i]]++] = i;
    // backward levels of the leading rows (the block's rows are solved first)
    nl2 = 0;
    for (int i = s - 1; i >= 0; --i) {
      int l = 0;
      for (int p = P->dpos[i] + 1; p < P->rp[i + 1]; ++p)
        if (P->ci[p] < s) l = std::max(l, lv[P->ci[p]] + 1);
      lv[i] = l;
      nl2 = std::max(nl2, l + 1);
    }
    P->nlev_lb = nl2;
    P->lbl_ptr.assign(nl2 + 1, 0);
    for (int i = 0; i < s; ++i) P->lbl_ptr[lv[i] + 1]++;
    for (int l = 0; l < nl2; ++l) P->lbl_ptr[l + 1] += P->lbl_ptr[l];
    std::vector<int> fill2(P->lbl_ptr.begin(), P->lbl_ptr.end() - 1);
    P->lbl_rows.assign(std::max(s, 1), 0);
    for (int i = 0; i < s; ++i) P->lbl_rows[fill2[lv[i]]++] = i;
  }
  // positions in old indices
  auto pos_of = [&](int orow, int ocol) {
    const int i = iperm[orow], c = iperm[ocol];
    const auto b = P->ci.begin() + P->rp[i], e = P->ci.begin() + P->rp[i + 1];
    const auto it = std::lower_bound(b, e, c);
    return (it != e && *it == c) ? (int)(it - P->ci.begin()) : -1;
  };
  P->pos_row.resize(P->nnz);
  P->pos_col.resize(P->nnz);
  for (int i = 0; i < N; ++i)
    for (int p = P->rp[i]; p < P->rp[i + 1]; ++p) {
      P->pos_row[p] = perm[i];
      P->pos_col[p] = perm[P->ci[p]];
    }
  // Jacobian terms in the reference's order of accumulation (every
  // destruction term, then every creation term, each in step order:
  // reactor_nvector_serial.cpp:146-172, 655-669)
  std::vector<std::vector<uint32_t>> dl(P->nnz), cl(P->nnz);
  for (int s = 0; s < M.nstep; ++s)
    for (int k = M.reac_ptr[s]; k < M.reac_ptr[s + 1]; ++k) {
      const int c = M.reac_sp[k];
      for (int q = M.reac_ptr[s]; q < M.reac_ptr[s + 1]; ++q)
        dl[pos_of(M.reac_sp[q], c)].push_back((uint32_t)s | (1u << 31));
      for (int q = M.prod_ptr[s]; q < M.prod_ptr[s + 1]; ++q)
        cl[pos_of(M.prod_sp[q], c)].push_back((uint32_t)s);
    }
  P->jt_ptr.assign(P->nnz + 1, 0);
  P->jt.clear();
  for (int p = 0; p < P->nnz; ++p) {
    P->jt.insert(P->jt.end(), dl[p].begin(), dl[p].end());
    P->jt.insert(P->jt.end(), cl[p].begin(), cl[p].end());
    P->jt_ptr[p + 1] = (int)P->jt.size();
  }
  if (P->jt.empty()) P->jt.push_back(0);
  // species rows of each old species column, ascending old row
  P->cp_ptr.assign(nsp + 1, 0);
  P->cp_pos.clear();
  for (int j = 0; j < nsp; ++j) {
    for (int r = 0; r < nsp; ++r) {
      const int p = (rows[r].count(j) != 0) ? pos_of(r, j) : -1;
      if (p >= 0) P->cp_pos.push_back(p);
    }
    P->cp_ptr[j + 1] = (int)P->cp_pos.size();
  }
  P->trow_pos.resize(nsp);
  for (int j = 0; j < nsp; ++j) P->trow_pos[j] = pos_of(nsp, j);
  P->tcol_pos.resize(N);
  for (int k = 0; k < N; ++k) P->tcol_pos[k] = pos_of(k, nsp);
  for (int j = 0; j < nsp; ++j)
    if (P->trow_pos[j] < 0 || P->tcol_pos[j] < 0) {
      *err = "temperature row or column missing from the pattern";
      return false;
    }
  return true;
}

}  // namespace zrk
