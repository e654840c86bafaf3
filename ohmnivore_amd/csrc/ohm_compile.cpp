// ohm_compile.cpp — MNA compiler: compiler::compile (src/compiler.rs:194-754).
//
// Produces the reference's CSR bit for bit: unknowns are the non-ground
// nodes in first-appearance order (:195-232) followed by V-source and
// inductor branches in component order (:236-252); stamps are emitted in
// component order (:306-568), then GMIN on every node diagonal (:571-573),
// then a zero for every C position (:577-579). The triplet list is ordered
// by (row, col) with ties kept in emission order and duplicates summed in
// that order — CsrMatrix::from_triplets (src/sparse.rs:52-100). Here that
// stable order comes from a counting sort by row followed by a stable sort of
// each row by column, which is O(nnz) for the short rows of circuit
// matrices.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ohm_internal.h"

namespace {

constexpr double kVt300 = 0.02585;  // src/compiler.rs:40
constexpr double kGmin = 1e-12;     // src/compiler.rs:44

struct Trip {
  std::vector<int64_t> r, c;
  std::vector<double> v;
  void push(int64_t rr, int64_t cc, double vv) {
    r.push_back(rr);
    c.push_back(cc);
    v.push_back(vv);
  }
};

bool ground_name(const char *s) {
  if (s[0] == '0' && s[1] == 0) return true;
  return std::tolower((unsigned char)s[0]) == 'g' && std::tolower((unsigned char)s[1]) == 'n' &&
         std::tolower((unsigned char)s[2]) == 'd' && s[3] == 0;
}

// Stable (row, col) ordering with in-order duplicate summation.
void to_csr(int64_t n, const Trip &t, std::vector<uint32_t> &rowptr, std::vector<uint32_t> &col,
            std::vector<double> &val) {
  const size_t m = t.r.size();
  std::vector<uint32_t> cnt(n + 1, 0);
  for (size_t k = 0; k < m; k++) cnt[t.r[k] + 1]++;
  for (int64_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  std::vector<uint32_t> order(m);
  {
    std::vector<uint32_t> pos(cnt.begin(), cnt.end() - 1);
    for (size_t k = 0; k < m; k++) order[pos[t.r[k]]++] = (uint32_t)k;
  }
  rowptr.assign(n + 1, 0);
  col.clear();
  val.clear();
  col.reserve(m);
  val.reserve(m);
  for (int64_t i = 0; i < n; i++) {
    uint32_t *b = order.data() + cnt[i], *e = order.data() + cnt[i + 1];
    std::stable_sort(b, e, [&](uint32_t x, uint32_t y) { return t.c[x] < t.c[y]; });
    for (uint32_t *p = b; p < e;) {
      int64_t c = t.c[*p];
      double s = t.v[*p];
      uint32_t *q = p + 1;
      while (q < e && t.c[*q] == c) s += t.v[*q++];
      col.push_back((uint32_t)c);
      val.push_back(s);
      p = q;
    }
    rowptr[i + 1] = (uint32_t)col.size();
  }
}

uint32_t slot(const std::vector<uint32_t> &rp, const std::vector<uint32_t> &cl, int64_t r,
              int64_t c) {
  if (r < 0 || c < 0) return OHM_SENTINEL;
  auto b = cl.begin() + rp[r], e = cl.begin() + rp[r + 1];
  auto it = std::lower_bound(b, e, (uint32_t)c);
  return (it != e && *it == (uint32_t)c) ? (uint32_t)(it - cl.begin()) : OHM_SENTINEL;
}

template <typename T>
T *dup(const std::vector<T> &v) {
  T *p = (T *)std::malloc(sizeof(T) * (v.size() ? v.size() : 1));
  if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

}  // namespace

using ohm::fail;

extern "C" ohm_status ohm_compile(const ohm_circuit *c, ohm_mna *out) {
  if (!c || !out) return fail(OHM_ERR_INVALID, "null argument");
  std::memset(out, 0, sizeof *out);
  const int32_t nc = c->n_components;
  std::vector<char> gnd(c->n_node_names);
  std::vector<int32_t> idx(c->n_node_names, -1), order;
  for (int32_t i = 0; i < c->n_node_names; i++) gnd[i] = ground_name(c->node_names[i]);
  for (int32_t k = 0; k < nc; k++) {
    const ohm_component &p = c->components[k];
    if (p.kind < OHM_COMP_R || p.kind > OHM_COMP_M)
      return fail(OHM_ERR_INVALID, "unknown component kind %d", p.kind);
    int nt = (p.kind == OHM_COMP_Q || p.kind == OHM_COMP_M) ? 3 : 2;
    for (int t = 0; t < nt; t++) {
      int32_t nm = p.node[t];
      if (nm < 0 || nm >= c->n_node_names)
        return fail(OHM_ERR_INVALID, "component %d references node %d outside the name table", k,
                    nm);
      if (!gnd[nm] && idx[nm] < 0) {
        idx[nm] = (int32_t)order.size();
        order.push_back(nm);
      }
    }
  }
  const int64_t n_nodes = (int64_t)order.size();
  std::vector<int32_t> branch_of(nc, -1), branch_comp;
  for (int32_t k = 0; k < nc; k++)
    if (c->components[k].kind == OHM_COMP_V || c->components[k].kind == OHM_COMP_L) {
      branch_of[k] = (int32_t)branch_comp.size();
      branch_comp.push_back(k);
    }
  const int64_t size = n_nodes + (int64_t)branch_comp.size();
  Trip gt, ct;
  gt.r.reserve(4 * (size_t)nc + n_nodes);
  std::vector<double> b_dc(size, 0.0), b_ac(2 * size, 0.0);
  struct PD { int64_t a, k; float is, nvt; };
  struct PQ { int64_t n[3]; float pol, is, bf, br, nf, nr; };
  struct PM { int64_t n[3]; float pol, vto, kp, lam; };
  std::vector<PD> pd;
  std::vector<PQ> pq;
  std::vector<PM> pm;
  auto name = [&](int32_t k) { return c->component_names ? c->component_names[k] : "?"; };
  auto mname = [&](int32_t k) {
    return (c->model_names && c->model_names[k]) ? c->model_names[k] : "";
  };
  for (int32_t k = 0; k < nc; k++) {
    const ohm_component &p = c->components[k];
    auto node = [&](int t) -> int64_t { return gnd[p.node[t]] ? -1 : idx[p.node[t]]; };
    switch (p.kind) {
      case OHM_COMP_R: {
        if (p.value == 0.0)
          return fail(OHM_ERR_COMPILE, "Compile error: Resistor %s has zero resistance", name(k));
        double g = 1.0 / p.value;
        int64_t i = node(0), j = node(1);
        if (i >= 0) gt.push(i, i, g);
        if (j >= 0) gt.push(j, j, g);
        if (i >= 0 && j >= 0) {
          gt.push(i, j, -g);
          gt.push(j, i, -g);
        }
        break;
      }
      case OHM_COMP_C: {
        if (p.value < 0.0)
          return fail(OHM_ERR_COMPILE, "Compile error: Capacitor %s has negative capacitance",
                      name(k));
        int64_t i = node(0), j = node(1);
        if (i >= 0) ct.push(i, i, p.value);
        if (j >= 0) ct.push(j, j, p.value);
        if (i >= 0 && j >= 0) {
          ct.push(i, j, -p.value);
          ct.push(j, i, -p.value);
        }
        break;
      }
      case OHM_COMP_L:
      case OHM_COMP_V: {
        int64_t bk = n_nodes + branch_of[k];
        int64_t i = node(0), j = node(1);
        if (i >= 0) {
          gt.push(i, bk, 1.0);
          gt.push(bk, i, 1.0);
        }
        if (j >= 0) {
          gt.push(j, bk, -1.0);
          gt.push(bk, j, -1.0);
        }
        if (p.kind == OHM_COMP_L) {
          ct.push(bk, bk, -p.value);
        } else {
          if (p.flags & OHM_COMP_HAS_DC) b_dc[bk] = p.value;
          if (p.flags & OHM_COMP_HAS_AC) {
            double ph = p.ac_phase_deg * (M_PI / 180.0);
            b_ac[2 * bk] = p.ac_mag * std::cos(ph);
            b_ac[2 * bk + 1] = p.ac_mag * std::sin(ph);
          }
        }
        break;
      }
      case OHM_COMP_D: {
        if (p.model < 0 || p.model >= c->n_diode_models)
          return fail(OHM_ERR_COMPILE, "Compile error: Diode %s references undefined model '%s'",
                      name(k), mname(k));
        const ohm_diode_model &m = c->diode_models[p.model];
        int64_t i = node(0), j = node(1);
        if (i >= 0) gt.push(i, i, 0.0);
        if (j >= 0) gt.push(j, j, 0.0);
        if (i >= 0 && j >= 0) {
          gt.push(i, j, 0.0);
          gt.push(j, i, 0.0);
        }
        pd.push_back({i, j, (float)m.is, (float)(m.n * kVt300)});
        break;
      }
      case OHM_COMP_I: {
        int64_t i = node(0), j = node(1);
        if (p.flags & OHM_COMP_HAS_DC) {
          if (i >= 0) b_dc[i] -= p.value;
          if (j >= 0) b_dc[j] += p.value;
        }
        if (p.flags & OHM_COMP_HAS_AC) {
          double ph = p.ac_phase_deg * (M_PI / 180.0);
          double re = p.ac_mag * std::cos(ph), im = p.ac_mag * std::sin(ph);
          if (i >= 0) {
            b_ac[2 * i] -= re;
            b_ac[2 * i + 1] -= im;
          }
          if (j >= 0) {
            b_ac[2 * j] += re;
            b_ac[2 * j + 1] += im;
          }
        }
        break;
      }
      case OHM_COMP_Q: {
        if (p.model < 0 || p.model >= c->n_bjt_models)
          return fail(OHM_ERR_COMPILE, "Compile error: BJT %s references undefined model '%s'",
                      name(k), mname(k));
        const ohm_bjt_model &m = c->bjt_models[p.model];
        PQ q{{node(0), node(1), node(2)}, m.is_npn ? 1.0f : -1.0f, (float)m.is, (float)m.bf,
             (float)m.br, (float)(m.nf * kVt300), (float)(m.nr * kVt300)};
        for (int r = 0; r < 3; r++)
          for (int cc = 0; cc < 3; cc++)
            if (q.n[r] >= 0 && q.n[cc] >= 0) gt.push(q.n[r], q.n[cc], 0.0);
        pq.push_back(q);
        break;
      }
      case OHM_COMP_M: {
        if (p.model < 0 || p.model >= c->n_mosfet_models)
          return fail(OHM_ERR_COMPILE, "Compile error: MOSFET %s references undefined model '%s'",
                      name(k), mname(k));
        const ohm_mosfet_model &m = c->mosfet_models[p.model];
        PM q{{node(0), node(1), node(2)}, m.is_nmos ? 1.0f : -1.0f, (float)std::fabs(m.vto),
             (float)m.kp, (float)m.lambda};
        const int64_t rows[2] = {q.n[0], q.n[2]};
        for (int r = 0; r < 2; r++)
          for (int cc = 0; cc < 3; cc++)
            if (rows[r] >= 0 && q.n[cc] >= 0) gt.push(rows[r], q.n[cc], 0.0);
        pm.push_back(q);
        break;
      }
    }
  }
  for (int64_t i = 0; i < n_nodes; i++) gt.push(i, i, kGmin);
  for (size_t t = 0; t < ct.r.size(); t++) gt.push(ct.r[t], ct.c[t], 0.0);

  std::vector<uint32_t> grp, gcl, crp, ccl;
  std::vector<double> gv, cv;
  to_csr(size, gt, grp, gcl, gv);
  to_csr(size, ct, crp, ccl, cv);

  std::vector<char> fixed(n_nodes, 0);
  for (int32_t k = 0; k < nc; k++) {
    const ohm_component &p = c->components[k];
    if (p.kind != OHM_COMP_V) continue;
    for (int t = 0; t < 2; t++)
      if (!gnd[p.node[t]]) fixed[idx[p.node[t]]] = 1;
  }
  auto u32 = [](int64_t i) { return i >= 0 ? (uint32_t)i : OHM_SENTINEL; };
  std::vector<ohm_diode_desc> dd(pd.size());
  for (size_t k = 0; k < pd.size(); k++) {
    ohm_diode_desc &d = dd[k];
    std::memset(&d, 0, sizeof d);
    int64_t a = pd[k].a, ca = pd[k].k;
    d.anode_idx = u32(a);
    d.cathode_idx = u32(ca);
    d.is_val = pd[k].is;
    d.n_vt = pd[k].nvt;
    d.g_row_col[0] = slot(grp, gcl, a, a);
    d.g_row_col[1] = slot(grp, gcl, ca, ca);
    d.g_row_col[2] = (a >= 0 && ca >= 0) ? slot(grp, gcl, a, ca) : OHM_SENTINEL;
    d.g_row_col[3] = (a >= 0 && ca >= 0) ? slot(grp, gcl, ca, a) : OHM_SENTINEL;
    d.b_idx[0] = u32(a);
    d.b_idx[1] = u32(ca);
  }
  std::vector<ohm_bjt_desc> qd(pq.size());
  for (size_t k = 0; k < pq.size(); k++) {
    ohm_bjt_desc &d = qd[k];
    const PQ &q = pq[k];
    std::memset(&d, 0, sizeof d);
    d.collector_idx = u32(q.n[0]);
    d.base_idx = u32(q.n[1]);
    d.emitter_idx = u32(q.n[2]);
    d.polarity = q.pol;
    d.is_val = q.is;
    d.bf = q.bf;
    d.br = q.br;
    d.nf_vt = q.nf;
    d.nr_vt = q.nr;
    for (int r = 0; r < 3; r++)
      for (int cc = 0; cc < 3; cc++) d.g_row_col[r * 3 + cc] = slot(grp, gcl, q.n[r], q.n[cc]);
    for (int t = 0; t < 3; t++) {
      d.node_fixed[t] = q.n[t] >= 0 ? (uint32_t)fixed[q.n[t]] : 0u;
      d.b_idx[t] = u32(q.n[t]);
    }
  }
  std::vector<ohm_mosfet_desc> md(pm.size());
  for (size_t k = 0; k < pm.size(); k++) {
    ohm_mosfet_desc &d = md[k];
    const PM &q = pm[k];
    std::memset(&d, 0, sizeof d);
    d.drain_idx = u32(q.n[0]);
    d.gate_idx = u32(q.n[1]);
    d.source_idx = u32(q.n[2]);
    d.polarity = q.pol;
    d.vto = q.vto;
    d.kp = q.kp;
    d.lambda = q.lam;
    const int64_t rows[2] = {q.n[0], q.n[2]};
    for (int r = 0; r < 2; r++)
      for (int cc = 0; cc < 3; cc++) d.g_row_col[r * 3 + cc] = slot(grp, gcl, rows[r], q.n[cc]);
    uint32_t f = 0;
    for (int t = 0; t < 3; t++)
      if (q.n[t] >= 0 && fixed[q.n[t]]) f |= 1u << t;
    d.node_fixed_packed = f;
    d.b_idx[0] = u32(q.n[0]);
    d.b_idx[1] = u32(q.n[2]);
  }
  out->size = size;
  out->n_nodes = n_nodes;
  out->n_branches = (int64_t)branch_comp.size();
  out->g_nnz = (int64_t)gcl.size();
  out->c_nnz = (int64_t)ccl.size();
  out->g_rowptr = dup(grp);
  out->g_col = dup(gcl);
  out->g_val = dup(gv);
  out->c_rowptr = dup(crp);
  out->c_col = dup(ccl);
  out->c_val = dup(cv);
  out->b_dc = dup(b_dc);
  out->b_ac = dup(b_ac);
  out->node_order = dup(order);
  out->branch_component = dup(branch_comp);
  out->n_diodes = (int64_t)dd.size();
  out->n_bjts = (int64_t)qd.size();
  out->n_mosfets = (int64_t)md.size();
  out->diodes = dup(dd);
  out->bjts = dup(qd);
  out->mosfets = dup(md);
  return OHM_OK;
}

extern "C" void ohm_mna_free(ohm_mna *m) {
  if (!m) return;
  void *ptrs[] = {m->g_rowptr, m->g_col,   m->g_val,      m->c_rowptr,        m->c_col,
                  m->c_val,    m->b_dc,    m->b_ac,       m->node_order,      m->branch_component,
                  m->diodes,   m->bjts,    m->mosfets};
  for (void *p : ptrs) std::free(p);
  std::memset(m, 0, sizeof *m);
}
