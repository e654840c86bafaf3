/*
 * ohm_oracle.c — CPU oracle (plain C) for the ohmnivore solve path.
 *
 * TEST INFRASTRUCTURE ONLY: a restatement of the reference's algorithms used
 * as the reference answer in tests/ and as the timed CPU baseline in bench.py. The
 * product (ohmnivore_amd/) never links or calls it.
 *
 * All citations are to files of the reference repository
 * (winstanley-industries/ohmnivore).
 */
#include "ohm_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread char g_err[512];

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char *oo_last_error(void) { return g_err; }
void oo_free(void *p) { free(p); }

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
    abort();
  }
  return p;
}
static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

/* ------------------------------------------------------------------------ */
/* Complex arithmetic with num-complex 0.4 semantics (naive division).       */
/* ------------------------------------------------------------------------ */
typedef struct { double re, im; } cplx;
static inline cplx cadd(cplx a, cplx b) { return (cplx){a.re + b.re, a.im + b.im}; }
static inline cplx csub(cplx a, cplx b) { return (cplx){a.re - b.re, a.im - b.im}; }
static inline cplx cmul(cplx a, cplx b) {
  return (cplx){a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
static inline cplx cdiv(cplx a, cplx b) {
  double d = b.re * b.re + b.im * b.im;
  return (cplx){(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
static inline double cnorm(cplx a) { return hypot(a.re, a.im); }
static inline cplx cconj(cplx a) { return (cplx){a.re, -a.im}; }

/* ------------------------------------------------------------------------ */
/* CsrMatrix::from_triplets — src/sparse.rs:52-100.                          */
/* Stable sort of triplet indices by (row, col) (`sort_by_key`, :59), then   */
/* duplicates summed in insertion order (:72-73).                            */
/* ------------------------------------------------------------------------ */
static const int64_t *g_sort_rows, *g_sort_cols;
static int trip_cmp(const void *pa, const void *pb) {
  int64_t a = *(const int64_t *)pa, b = *(const int64_t *)pb;
  if (g_sort_rows[a] != g_sort_rows[b]) return g_sort_rows[a] < g_sort_rows[b] ? -1 : 1;
  if (g_sort_cols[a] != g_sort_cols[b]) return g_sort_cols[a] < g_sort_cols[b] ? -1 : 1;
  return a < b ? -1 : (a > b); /* index tie-break == stable sort */
}

int64_t oo_from_triplets(int64_t nrows, int64_t ntrip, const int64_t *rows, const int64_t *cols,
                         const double *vals, int complex_vals, int64_t *rowptr, int64_t *out_cols,
                         double *out_vals) {
  for (int64_t i = 0; i <= nrows; i++) rowptr[i] = 0;
  if (ntrip == 0) return 0;
  int64_t *idx = (int64_t *)xmalloc(sizeof(int64_t) * ntrip);
  for (int64_t i = 0; i < ntrip; i++) idx[i] = i;
  g_sort_rows = rows;
  g_sort_cols = cols;
  qsort(idx, ntrip, sizeof(int64_t), trip_cmp);
  int w = complex_vals ? 2 : 1;
  int64_t nnz = 0;
  int64_t first = idx[0];
  int64_t cur_row = rows[first], cur_col = cols[first];
  double cur_re = vals[w * first], cur_im = complex_vals ? vals[w * first + 1] : 0.0;
  for (int64_t k = 1; k < ntrip; k++) {
    int64_t t = idx[k];
    if (rows[t] == cur_row && cols[t] == cur_col) {
      cur_re += vals[w * t];
      if (complex_vals) cur_im += vals[w * t + 1];
    } else {
      out_vals[w * nnz] = cur_re;
      if (complex_vals) out_vals[w * nnz + 1] = cur_im;
      out_cols[nnz] = cur_col;
      rowptr[cur_row + 1] += 1;
      nnz++;
      cur_row = rows[t];
      cur_col = cols[t];
      cur_re = vals[w * t];
      cur_im = complex_vals ? vals[w * t + 1] : 0.0;
    }
  }
  out_vals[w * nnz] = cur_re;
  if (complex_vals) out_vals[w * nnz + 1] = cur_im;
  out_cols[nnz] = cur_col;
  rowptr[cur_row + 1] += 1;
  nnz++;
  for (int64_t i = 1; i <= nrows; i++) rowptr[i] += rowptr[i - 1];
  free(idx);
  return nnz;
}

/* ------------------------------------------------------------------------ */
/* compiler::compile — src/compiler.rs:194-754.                              */
/* ------------------------------------------------------------------------ */
#define VT_300K 0.02585 /* src/compiler.rs:40 */
#define GMIN 1e-12      /* src/compiler.rs:44 */

typedef struct {
  int64_t n, cap;
  int64_t *r, *c;
  double *v;
} tripvec;

static void tv_push(tripvec *t, int64_t r, int64_t c, double v) {
  if (t->n == t->cap) {
    t->cap = t->cap ? 2 * t->cap : 1024;
    t->r = (int64_t *)realloc(t->r, sizeof(int64_t) * t->cap);
    t->c = (int64_t *)realloc(t->c, sizeof(int64_t) * t->cap);
    t->v = (double *)realloc(t->v, sizeof(double) * t->cap);
    if (!t->r || !t->c || !t->v) abort();
  }
  t->r[t->n] = r;
  t->c[t->n] = c;
  t->v[t->n] = v;
  t->n++;
}
static void tv_free(tripvec *t) {
  free(t->r);
  free(t->c);
  free(t->v);
}

/* is_ground — src/compiler.rs:178-180 */
static int is_ground_name(const char *s) {
  if (strcmp(s, "0") == 0) return 1;
  return (tolower((unsigned char)s[0]) == 'g' && tolower((unsigned char)s[1]) == 'n' &&
          tolower((unsigned char)s[2]) == 'd' && s[3] == 0);
}

/* value_index — src/sparse.rs:24-30 (linear search within the row). */
static uint32_t value_index(const int64_t *rowptr, const int64_t *col, int64_t r, int64_t c) {
  for (int64_t k = rowptr[r]; k < rowptr[r + 1]; k++)
    if (col[k] == c) return (uint32_t)k;
  return OHM_SENTINEL;
}

typedef struct { int64_t a, c; float is_val, n_vt; } pend_diode;
typedef struct { int64_t n[3]; float pol, is_val, bf, br, nf_vt, nr_vt; } pend_bjt;
typedef struct { int64_t n[3]; float pol, vto, kp, lambda; } pend_mos;

int oo_compile(const ohm_circuit *c, ohm_mna *out) {
  memset(out, 0, sizeof *out);
  /* Step 1: node registration in first-appearance order (:195-232). */
  int32_t *name_to_idx = (int32_t *)xmalloc(sizeof(int32_t) * (c->n_node_names + 1));
  char *grounded = (char *)xcalloc(c->n_node_names + 1, 1);
  for (int32_t i = 0; i < c->n_node_names; i++) {
    name_to_idx[i] = -1;
    grounded[i] = (char)is_ground_name(c->node_names[i]);
  }
  int32_t *node_order = (int32_t *)xmalloc(sizeof(int32_t) * (c->n_node_names + 1));
  int64_t n_nodes = 0;
  for (int32_t k = 0; k < c->n_components; k++) {
    const ohm_component *p = &c->components[k];
    int nt = (p->kind == OHM_COMP_Q || p->kind == OHM_COMP_M) ? 3 : 2;
    for (int t = 0; t < nt; t++) {
      int32_t nm = p->node[t];
      if (grounded[nm]) continue;
      if (name_to_idx[nm] < 0) {
        name_to_idx[nm] = (int32_t)n_nodes;
        node_order[n_nodes++] = nm;
      }
    }
  }
  /* Step 2: branches = V sources and inductors in component order (:236-252). */
  int32_t *branch_of = (int32_t *)xmalloc(sizeof(int32_t) * (c->n_components + 1));
  int32_t *branch_comp = (int32_t *)xmalloc(sizeof(int32_t) * (c->n_components + 1));
  int64_t n_br = 0;
  for (int32_t k = 0; k < c->n_components; k++) {
    branch_of[k] = -1;
    int kd = c->components[k].kind;
    if (kd == OHM_COMP_V || kd == OHM_COMP_L) {
      branch_of[k] = (int32_t)n_br;
      branch_comp[n_br++] = k;
    }
  }
  int64_t size = n_nodes + n_br;
  tripvec gt = {0}, ct = {0};
  double *b_dc = (double *)xcalloc(size + 1, sizeof(double));
  double *b_ac = (double *)xcalloc(2 * size + 2, sizeof(double));
  pend_diode *pd = (pend_diode *)xmalloc(sizeof(pend_diode) * (c->n_components + 1));
  pend_bjt *pb = (pend_bjt *)xmalloc(sizeof(pend_bjt) * (c->n_components + 1));
  pend_mos *pm = (pend_mos *)xmalloc(sizeof(pend_mos) * (c->n_components + 1));
  int64_t nd = 0, nq = 0, nmos = 0;
  int rc = OHM_OK;
#define NODE(t) (grounded[p->node[t]] ? (int64_t)-1 : (int64_t)name_to_idx[p->node[t]])
  /* Step 3: stamps in component order (:306-568). */
  for (int32_t k = 0; k < c->n_components && rc == OHM_OK; k++) {
    const ohm_component *p = &c->components[k];
    const char *nm = c->component_names ? c->component_names[k] : "?";
    switch (p->kind) {
      case OHM_COMP_R: { /* :308-329 */
        double g = 1.0 / p->value;
        if (p->value == 0.0) {
          rc = fail(OHM_ERR_COMPILE, "Compile error: Resistor %s has zero resistance", nm);
          break;
        }
        int64_t i = NODE(0), j = NODE(1);
        if (i >= 0) tv_push(&gt, i, i, g);
        if (j >= 0) tv_push(&gt, j, j, g);
        if (i >= 0 && j >= 0) {
          tv_push(&gt, i, j, -g);
          tv_push(&gt, j, i, -g);
        }
        break;
      }
      case OHM_COMP_C: { /* :331-352 */
        double cap = p->value;
        if (cap < 0.0) {
          rc = fail(OHM_ERR_COMPILE, "Compile error: Capacitor %s has negative capacitance", nm);
          break;
        }
        int64_t i = NODE(0), j = NODE(1);
        if (i >= 0) tv_push(&ct, i, i, cap);
        if (j >= 0) tv_push(&ct, j, j, cap);
        if (i >= 0 && j >= 0) {
          tv_push(&ct, i, j, -cap);
          tv_push(&ct, j, i, -cap);
        }
        break;
      }
      case OHM_COMP_L:   /* :354-374 */
      case OHM_COMP_V: { /* :376-408 */
        int64_t bk = n_nodes + branch_of[k];
        int64_t i = NODE(0), j = NODE(1);
        if (i >= 0) {
          tv_push(&gt, i, bk, 1.0);
          tv_push(&gt, bk, i, 1.0);
        }
        if (j >= 0) {
          tv_push(&gt, j, bk, -1.0);
          tv_push(&gt, bk, j, -1.0);
        }
        if (p->kind == OHM_COMP_L) {
          tv_push(&ct, bk, bk, -p->value);
        } else {
          if (p->flags & OHM_COMP_HAS_DC) b_dc[bk] = p->value;
          if (p->flags & OHM_COMP_HAS_AC) {
            double ph = p->ac_phase_deg * (M_PI / 180.0); /* f64::to_radians */
            b_ac[2 * bk] = p->ac_mag * cos(ph);
            b_ac[2 * bk + 1] = p->ac_mag * sin(ph);
          }
        }
        break;
      }
      case OHM_COMP_D: { /* :412-448 */
        if (p->model < 0) {
          rc = fail(OHM_ERR_COMPILE, "Compile error: Diode %s references undefined model '%s'", nm,
                    c->model_names && c->model_names[k] ? c->model_names[k] : "");
          break;
        }
        const ohm_diode_model *m = &c->diode_models[p->model];
        int64_t i = NODE(0), j = NODE(1);
        if (i >= 0) tv_push(&gt, i, i, 0.0);
        if (j >= 0) tv_push(&gt, j, j, 0.0);
        if (i >= 0 && j >= 0) {
          tv_push(&gt, i, j, 0.0);
          tv_push(&gt, j, i, 0.0);
        }
        pd[nd++] = (pend_diode){i, j, (float)m->is, (float)(m->n * VT_300K)};
        break;
      }
      case OHM_COMP_I: { /* :450-479 */
        int64_t i = NODE(0), j = NODE(1);
        if (p->flags & OHM_COMP_HAS_DC) {
          if (i >= 0) b_dc[i] -= p->value;
          if (j >= 0) b_dc[j] += p->value;
        }
        if (p->flags & OHM_COMP_HAS_AC) {
          double ph = p->ac_phase_deg * (M_PI / 180.0);
          double re = p->ac_mag * cos(ph), im = p->ac_mag * sin(ph);
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
      case OHM_COMP_Q: { /* :482-522 */
        if (p->model < 0) {
          rc = fail(OHM_ERR_COMPILE, "Compile error: BJT %s references undefined model '%s'", nm,
                    c->model_names && c->model_names[k] ? c->model_names[k] : "");
          break;
        }
        const ohm_bjt_model *m = &c->bjt_models[p->model];
        int64_t ni[3] = {NODE(0), NODE(1), NODE(2)};
        for (int r = 0; r < 3; r++)
          for (int cc = 0; cc < 3; cc++)
            if (ni[r] >= 0 && ni[cc] >= 0) tv_push(&gt, ni[r], ni[cc], 0.0);
        pend_bjt q;
        for (int t = 0; t < 3; t++) q.n[t] = ni[t];
        q.pol = m->is_npn ? 1.0f : -1.0f;
        q.is_val = (float)m->is;
        q.bf = (float)m->bf;
        q.br = (float)m->br;
        q.nf_vt = (float)(m->nf * VT_300K);
        q.nr_vt = (float)(m->nr * VT_300K);
        pb[nq++] = q;
        break;
      }
      case OHM_COMP_M: { /* :526-566 */
        if (p->model < 0) {
          rc = fail(OHM_ERR_COMPILE, "Compile error: MOSFET %s references undefined model '%s'", nm,
                    c->model_names && c->model_names[k] ? c->model_names[k] : "");
          break;
        }
        const ohm_mosfet_model *m = &c->mosfet_models[p->model];
        int64_t ni[3] = {NODE(0), NODE(1), NODE(2)};
        int64_t rows[2] = {ni[0], ni[2]};
        for (int r = 0; r < 2; r++)
          for (int cc = 0; cc < 3; cc++)
            if (rows[r] >= 0 && ni[cc] >= 0) tv_push(&gt, rows[r], ni[cc], 0.0);
        pend_mos q;
        for (int t = 0; t < 3; t++) q.n[t] = ni[t];
        q.pol = m->is_nmos ? 1.0f : -1.0f;
        q.vto = (float)fabs(m->vto);
        q.kp = (float)m->kp;
        q.lambda = (float)m->lambda;
        pm[nmos++] = q;
        break;
      }
      default:
        rc = fail(OHM_ERR_INVALID, "unknown component kind %d", p->kind);
    }
  }
#undef NODE
  if (rc != OHM_OK) goto cleanup_err;
  /* GMIN on every node diagonal (:571-573), then C -> G zero-fill (:577-579). */
  for (int64_t i = 0; i < n_nodes; i++) tv_push(&gt, i, i, GMIN);
  for (int64_t t = 0; t < ct.n; t++) tv_push(&gt, ct.r[t], ct.c[t], 0.0);

  /* Step 4: CSR (:582-583). */
  {
    int64_t *g_rp = (int64_t *)xmalloc(sizeof(int64_t) * (size + 1));
    int64_t *g_cl = (int64_t *)xmalloc(sizeof(int64_t) * (gt.n + 1));
    double *g_v = (double *)xmalloc(sizeof(double) * (gt.n + 1));
    int64_t gnnz = oo_from_triplets(size, gt.n, gt.r, gt.c, gt.v, 0, g_rp, g_cl, g_v);
    int64_t *c_rp = (int64_t *)xmalloc(sizeof(int64_t) * (size + 1));
    int64_t *c_cl = (int64_t *)xmalloc(sizeof(int64_t) * (ct.n + 1));
    double *c_v = (double *)xmalloc(sizeof(double) * (ct.n + 1));
    int64_t cnnz = oo_from_triplets(size, ct.n, ct.r, ct.c, ct.v, 0, c_rp, c_cl, c_v);

    /* V-source constrained node set (:587-597). */
    char *fixed = (char *)xcalloc(n_nodes + 1, 1);
    for (int32_t k = 0; k < c->n_components; k++) {
      const ohm_component *p = &c->components[k];
      if (p->kind != OHM_COMP_V) continue;
      for (int t = 0; t < 2; t++)
        if (!grounded[p->node[t]]) fixed[name_to_idx[p->node[t]]] = 1;
    }
#define VI(r, cc) (((r) >= 0 && (cc) >= 0) ? value_index(g_rp, g_cl, (r), (cc)) : OHM_SENTINEL)
#define U32(i) ((i) >= 0 ? (uint32_t)(i) : OHM_SENTINEL)
    /* Step 5: diode descriptors (:602-646). */
    out->diodes = (ohm_diode_desc *)xcalloc(nd + 1, sizeof(ohm_diode_desc));
    for (int64_t k = 0; k < nd; k++) {
      ohm_diode_desc *d = &out->diodes[k];
      int64_t a = pd[k].a, cc = pd[k].c;
      d->anode_idx = U32(a);
      d->cathode_idx = U32(cc);
      d->is_val = pd[k].is_val;
      d->n_vt = pd[k].n_vt;
      d->g_row_col[0] = VI(a, a);
      d->g_row_col[1] = VI(cc, cc);
      d->g_row_col[2] = VI(a, cc);
      d->g_row_col[3] = VI(cc, a);
      d->b_idx[0] = U32(a);
      d->b_idx[1] = U32(cc);
    }
    /* Step 6: BJT descriptors (:649-692). */
    out->bjts = (ohm_bjt_desc *)xcalloc(nq + 1, sizeof(ohm_bjt_desc));
    for (int64_t k = 0; k < nq; k++) {
      ohm_bjt_desc *d = &out->bjts[k];
      const pend_bjt *q = &pb[k];
      d->collector_idx = U32(q->n[0]);
      d->base_idx = U32(q->n[1]);
      d->emitter_idx = U32(q->n[2]);
      d->polarity = q->pol;
      d->is_val = q->is_val;
      d->bf = q->bf;
      d->br = q->br;
      d->nf_vt = q->nf_vt;
      d->nr_vt = q->nr_vt;
      for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) d->g_row_col[r * 3 + cc] = VI(q->n[r], q->n[cc]);
      for (int t = 0; t < 3; t++) {
        d->node_fixed[t] = q->n[t] >= 0 ? (uint32_t)fixed[q->n[t]] : 0u;
        d->b_idx[t] = U32(q->n[t]);
      }
    }
    /* Step 7: MOSFET descriptors (:695-740). */
    out->mosfets = (ohm_mosfet_desc *)xcalloc(nmos + 1, sizeof(ohm_mosfet_desc));
    for (int64_t k = 0; k < nmos; k++) {
      ohm_mosfet_desc *d = &out->mosfets[k];
      const pend_mos *q = &pm[k];
      d->drain_idx = U32(q->n[0]);
      d->gate_idx = U32(q->n[1]);
      d->source_idx = U32(q->n[2]);
      d->polarity = q->pol;
      d->vto = q->vto;
      d->kp = q->kp;
      d->lambda = q->lambda;
      int64_t rows[2] = {q->n[0], q->n[2]};
      for (int r = 0; r < 2; r++)
        for (int cc = 0; cc < 3; cc++) d->g_row_col[r * 3 + cc] = VI(rows[r], q->n[cc]);
      uint32_t f = 0;
      for (int t = 0; t < 3; t++)
        if (q->n[t] >= 0 && fixed[q->n[t]]) f |= 1u << t;
      d->node_fixed_packed = f;
      d->b_idx[0] = U32(q->n[0]);
      d->b_idx[1] = U32(q->n[2]);
    }
#undef VI
#undef U32
    free(fixed);
    out->size = size;
    out->n_nodes = n_nodes;
    out->n_branches = n_br;
    out->g_nnz = gnnz;
    out->c_nnz = cnnz;
    out->g_rowptr = (uint32_t *)xmalloc(sizeof(uint32_t) * (size + 1));
    out->c_rowptr = (uint32_t *)xmalloc(sizeof(uint32_t) * (size + 1));
    for (int64_t i = 0; i <= size; i++) {
      out->g_rowptr[i] = (uint32_t)g_rp[i];
      out->c_rowptr[i] = (uint32_t)c_rp[i];
    }
    out->g_col = (uint32_t *)xmalloc(sizeof(uint32_t) * (gnnz + 1));
    out->g_val = (double *)xmalloc(sizeof(double) * (gnnz + 1));
    for (int64_t k = 0; k < gnnz; k++) {
      out->g_col[k] = (uint32_t)g_cl[k];
      out->g_val[k] = g_v[k];
    }
    out->c_col = (uint32_t *)xmalloc(sizeof(uint32_t) * (cnnz + 1));
    out->c_val = (double *)xmalloc(sizeof(double) * (cnnz + 1));
    for (int64_t k = 0; k < cnnz; k++) {
      out->c_col[k] = (uint32_t)c_cl[k];
      out->c_val[k] = c_v[k];
    }
    free(g_rp);
    free(g_cl);
    free(g_v);
    free(c_rp);
    free(c_cl);
    free(c_v);
  }
  out->b_dc = b_dc;
  out->b_ac = b_ac;
  out->node_order = node_order;
  out->branch_component = branch_comp;
  out->n_diodes = nd;
  out->n_bjts = nq;
  out->n_mosfets = nmos;
  free(name_to_idx);
  free(grounded);
  free(branch_of);
  free(pd);
  free(pb);
  free(pm);
  tv_free(&gt);
  tv_free(&ct);
  return OHM_OK;

cleanup_err:
  free(name_to_idx);
  free(grounded);
  free(node_order);
  free(branch_of);
  free(branch_comp);
  free(b_dc);
  free(b_ac);
  free(pd);
  free(pb);
  free(pm);
  tv_free(&gt);
  tv_free(&ct);
  return rc;
}

void oo_mna_free(ohm_mna *m) {
  free(m->g_rowptr);
  free(m->g_col);
  free(m->g_val);
  free(m->c_rowptr);
  free(m->c_col);
  free(m->c_val);
  free(m->b_dc);
  free(m->b_ac);
  free(m->node_order);
  free(m->branch_component);
  free(m->diodes);
  free(m->bjts);
  free(m->mosfets);
  memset(m, 0, sizeof *m);
}

/* ------------------------------------------------------------------------ */
/* SpMV — src/sparse.rs:116-143 (sequential row sums).                       */
/* ------------------------------------------------------------------------ */
void oo_spmv(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
             const double *x, double *y, int cplx_vals) {
  if (!cplx_vals) {
    for (int64_t r = 0; r < n; r++) {
      double s = 0.0;
      for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++) s += val[k] * x[col[k]];
      y[r] = s;
    }
  } else {
    const cplx *v = (const cplx *)val, *xx = (const cplx *)x;
    cplx *yy = (cplx *)y;
    for (int64_t r = 0; r < n; r++) {
      cplx s = {0, 0};
      for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++) s = cadd(s, cmul(v[k], xx[col[k]]));
      yy[r] = s;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* CpuSolver — src/solver/cpu.rs:28-170: dense augmented [A|b], forward       */
/* elimination with partial pivoting (singular if max pivot < 1e-15), back   */
/* substitution.                                                             */
/* ------------------------------------------------------------------------ */
int oo_dense_solve(int64_t nrows, int64_t ncols, int64_t nb, const uint32_t *rowptr,
                   const uint32_t *col, const double *val, const double *b, double *x,
                   int cplx_vals) {
  if (ncols != nrows || nb != nrows)
    return fail(OHM_ERR_SOLVE, "Solve error: dimension mismatch: matrix is %lldx%lld, rhs length is %lld",
                (long long)nrows, (long long)ncols, (long long)nb);
  int64_t n = nrows;
  if (n == 0) return OHM_OK;
  int64_t w = n + 1;
  if (!cplx_vals) {
    double **aug = (double **)xmalloc(sizeof(double *) * n);
    for (int64_t i = 0; i < n; i++) {
      aug[i] = (double *)xcalloc(w, sizeof(double));
      for (uint32_t k = rowptr[i]; k < rowptr[i + 1]; k++) aug[i][col[k]] = val[k]; /* to_dense */
      aug[i][n] = b[i];
    }
    int rc = OHM_OK;
    for (int64_t k = 0; k < n && rc == OHM_OK; k++) {
      double max_val = fabs(aug[k][k]);
      int64_t max_row = k;
      for (int64_t i = k + 1; i < n; i++) {
        double v = fabs(aug[i][k]);
        if (v > max_val) {
          max_val = v;
          max_row = i;
        }
      }
      if (max_val < 1e-15) {
        rc = fail(OHM_ERR_SOLVE, "Solve error: singular matrix");
        break;
      }
      if (max_row != k) {
        double *t = aug[k];
        aug[k] = aug[max_row];
        aug[max_row] = t;
      }
      double pivot = aug[k][k];
      const double *rk = aug[k];
      for (int64_t i = k + 1; i < n; i++) {
        double *ri = aug[i];
        double factor = ri[k] / pivot;
        ri[k] = 0.0;
        for (int64_t j = k + 1; j <= n; j++) ri[j] -= factor * rk[j];
      }
    }
    if (rc == OHM_OK) {
      for (int64_t i = n - 1; i >= 0; i--) {
        double sum = aug[i][n];
        for (int64_t j = i + 1; j < n; j++) sum -= aug[i][j] * x[j];
        x[i] = sum / aug[i][i];
      }
    }
    for (int64_t i = 0; i < n; i++) free(aug[i]);
    free(aug);
    return rc;
  } else {
    const cplx *v = (const cplx *)val, *bb = (const cplx *)b;
    cplx *xx = (cplx *)x;
    cplx **aug = (cplx **)xmalloc(sizeof(cplx *) * n);
    for (int64_t i = 0; i < n; i++) {
      aug[i] = (cplx *)xcalloc(w, sizeof(cplx));
      for (uint32_t k = rowptr[i]; k < rowptr[i + 1]; k++) aug[i][col[k]] = v[k];
      aug[i][n] = bb[i];
    }
    int rc = OHM_OK;
    for (int64_t k = 0; k < n && rc == OHM_OK; k++) {
      double max_val = cnorm(aug[k][k]);
      int64_t max_row = k;
      for (int64_t i = k + 1; i < n; i++) {
        double vv = cnorm(aug[i][k]);
        if (vv > max_val) {
          max_val = vv;
          max_row = i;
        }
      }
      if (max_val < 1e-15) {
        rc = fail(OHM_ERR_SOLVE, "Solve error: singular matrix");
        break;
      }
      if (max_row != k) {
        cplx *t = aug[k];
        aug[k] = aug[max_row];
        aug[max_row] = t;
      }
      cplx pivot = aug[k][k];
      for (int64_t i = k + 1; i < n; i++) {
        cplx factor = cdiv(aug[i][k], pivot);
        aug[i][k] = (cplx){0, 0};
        for (int64_t j = k + 1; j <= n; j++) aug[i][j] = csub(aug[i][j], cmul(factor, aug[k][j]));
      }
    }
    if (rc == OHM_OK) {
      for (int64_t i = n - 1; i >= 0; i--) {
        cplx sum = aug[i][n];
        for (int64_t j = i + 1; j < n; j++) sum = csub(sum, cmul(aug[i][j], xx[j]));
        xx[i] = cdiv(sum, aug[i][i]);
      }
    }
    for (int64_t i = 0; i < n; i++) free(aug[i]);
    free(aug);
    return rc;
  }
}

/* ------------------------------------------------------------------------ */
/* Sparse direct LU (stand-in for faer 0.24 sp_lu, src/solver/sparse_direct.rs:12-61). */
/* Reverse Cuthill-McKee on pattern(A + A^T) (George & Liu 1981), then the   */
/* banded LU with partial pivoting of LAPACK xGBTF2 and the solve of xGBTRS. */
/* ------------------------------------------------------------------------ */
static void rcm_order(int64_t n, const uint32_t *rowptr, const uint32_t *col, int64_t *perm) {
  /* symmetric adjacency without diagonal */
  int64_t *deg = (int64_t *)xcalloc(n + 1, sizeof(int64_t));
  for (int64_t r = 0; r < n; r++)
    for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++)
      if ((int64_t)col[k] != r) {
        deg[r]++;
        deg[col[k]]++;
      }
  int64_t *ap = (int64_t *)xmalloc(sizeof(int64_t) * (n + 1));
  ap[0] = 0;
  for (int64_t i = 0; i < n; i++) ap[i + 1] = ap[i] + deg[i];
  int64_t *adj = (int64_t *)xmalloc(sizeof(int64_t) * (ap[n] + 1));
  int64_t *fill = (int64_t *)xmalloc(sizeof(int64_t) * (n + 1));
  for (int64_t i = 0; i < n; i++) fill[i] = ap[i];
  for (int64_t r = 0; r < n; r++)
    for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++)
      if ((int64_t)col[k] != r) {
        adj[fill[r]++] = col[k];
        adj[fill[col[k]]++] = r;
      }
  /* dedup each adjacency list and recompute degree */
  for (int64_t i = 0; i < n; i++) {
    int64_t *a = adj + ap[i];
    int64_t m = ap[i + 1] - ap[i];
    for (int64_t x = 1; x < m; x++) { /* insertion sort (lists are short) */
      int64_t v = a[x], y = x - 1;
      while (y >= 0 && a[y] > v) {
        a[y + 1] = a[y];
        y--;
      }
      a[y + 1] = v;
    }
    int64_t u = 0;
    for (int64_t x = 0; x < m; x++)
      if (u == 0 || a[u - 1] != a[x]) a[u++] = a[x];
    deg[i] = u;
  }
  char *vis = (char *)xcalloc(n + 1, 1);
  int64_t *level = (int64_t *)xmalloc(sizeof(int64_t) * (n + 1));
  int64_t *order = (int64_t *)xmalloc(sizeof(int64_t) * (n + 1));
  int64_t *tmp = (int64_t *)xmalloc(sizeof(int64_t) * (n + 1));
  int64_t pos = 0;
  for (;;) {
    /* lowest-degree unvisited node starts a component */
    int64_t start = -1;
    for (int64_t i = 0; i < n; i++)
      if (!vis[i] && (start < 0 || deg[i] < deg[start])) start = i;
    if (start < 0) break;
    /* pseudo-peripheral node: repeat BFS from the min-degree node of the last level */
    int64_t ecc = -1;
    for (int it = 0; it < 8; it++) {
      int64_t head = 0, tail = 0, maxl = 0;
      tmp[tail++] = start;
      vis[start] = 2;
      level[start] = 0;
      while (head < tail) {
        int64_t u = tmp[head++];
        for (int64_t e = ap[u]; e < ap[u] + deg[u]; e++) {
          int64_t v = adj[e];
          if (vis[v] == 0) {
            vis[v] = 2;
            level[v] = level[u] + 1;
            if (level[v] > maxl) maxl = level[v];
            tmp[tail++] = v;
          }
        }
      }
      int64_t best = start;
      for (int64_t q = 0; q < tail; q++) {
        int64_t u = tmp[q];
        if (level[u] == maxl && deg[u] < deg[best]) best = u;
        if (level[u] == maxl && best == start) best = u;
      }
      for (int64_t q = 0; q < tail; q++) vis[tmp[q]] = 0;
      if (maxl <= ecc) break;
      ecc = maxl;
      start = best;
    }
    /* Cuthill-McKee BFS, neighbours by ascending degree */
    int64_t head = pos;
    order[pos++] = start;
    vis[start] = 1;
    while (head < pos) {
      int64_t u = order[head++];
      int64_t m = 0;
      for (int64_t e = ap[u]; e < ap[u] + deg[u]; e++)
        if (!vis[adj[e]]) tmp[m++] = adj[e];
      for (int64_t x = 1; x < m; x++) {
        int64_t v = tmp[x], y = x - 1;
        while (y >= 0 && (deg[tmp[y]] > deg[v] || (deg[tmp[y]] == deg[v] && tmp[y] > v))) {
          tmp[y + 1] = tmp[y];
          y--;
        }
        tmp[y + 1] = v;
      }
      for (int64_t x = 0; x < m; x++) {
        vis[tmp[x]] = 1;
        order[pos++] = tmp[x];
      }
    }
  }
  for (int64_t i = 0; i < n; i++) perm[i] = order[n - 1 - i]; /* reverse */
  free(deg);
  free(ap);
  free(adj);
  free(fill);
  free(vis);
  free(level);
  free(order);
  free(tmp);
}

/* Numeric banded LU with partial pivoting under a given symmetric ordering
 * (perm: new position -> old row), then the two triangular solves. */
static int band_lu_solve(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                         const double *b, double *x, int cplx_vals, const int64_t *perm) {
  int w = cplx_vals ? 2 : 1;
  for (int64_t k = 0; k < (int64_t)rowptr[n] * w; k++)
    if (!isfinite(val[k])) return fail(OHM_ERR_SOLVE, "Solve error: sparse LU input contains NaN/Inf");
  int64_t *iperm = (int64_t *)xmalloc(sizeof(int64_t) * n);
  for (int64_t i = 0; i < n; i++) iperm[perm[i]] = i;
  int64_t kl = 0, ku = 0;
  for (int64_t r = 0; r < n; r++)
    for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
      int64_t i = iperm[r], j = iperm[col[k]];
      if (i - j > kl) kl = i - j;
      if (j - i > ku) ku = j - i;
    }
  int64_t kv = kl + ku, ldab = 2 * kl + ku + 1;
  size_t ab_n = (size_t)ldab * (size_t)n;
  double *ab = (double *)xcalloc(ab_n * w, sizeof(double));
#define AB(i, j) ab[((size_t)(kv + (i) - (j)) + (size_t)(j) * (size_t)ldab) * w]
  for (int64_t r = 0; r < n; r++)
    for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
      int64_t i = iperm[r], j = iperm[col[k]];
      double *p = &AB(i, j);
      p[0] += val[k * w];
      if (cplx_vals) p[1] += val[k * w + 1];
    }
  int64_t *ipiv = (int64_t *)xmalloc(sizeof(int64_t) * n);
  int singular = 0;
  int64_t ju = 0;
  for (int64_t j = 0; j < n; j++) {
    int64_t km = kl < n - 1 - j ? kl : n - 1 - j;
    int64_t jp = 0;
    double best = -1.0;
    for (int64_t i = 0; i <= km; i++) {
      double *p = &AB(j + i, j);
      double a = cplx_vals ? cnorm((cplx){p[0], p[1]}) : fabs(p[0]);
      if (a > best) {
        best = a;
        jp = i;
      }
    }
    ipiv[j] = j + jp;
    if (best == 0.0) {
      singular = 1;
      continue;
    }
    int64_t jmax = j + ku + jp < n - 1 ? j + ku + jp : n - 1;
    if (jmax > ju) ju = jmax;
    if (jp != 0)
      for (int64_t c = j; c <= ju; c++) {
        double *p = &AB(j, c), *q = &AB(j + jp, c);
        for (int t = 0; t < w; t++) {
          double s = p[t];
          p[t] = q[t];
          q[t] = s;
        }
      }
    if (km > 0) {
      double *pv = &AB(j, j);
      if (!cplx_vals) {
        double rp = 1.0 / pv[0];
        for (int64_t i = 1; i <= km; i++) AB(j + i, j) *= rp;
        for (int64_t c = j + 1; c <= ju; c++) {
          double u = AB(j, c);
          if (u == 0.0) continue;
          for (int64_t i = 1; i <= km; i++) AB(j + i, c) -= AB(j + i, j) * u;
        }
      } else {
        cplx piv = {pv[0], pv[1]};
        for (int64_t i = 1; i <= km; i++) {
          double *q = &AB(j + i, j);
          cplx l = cdiv((cplx){q[0], q[1]}, piv);
          q[0] = l.re;
          q[1] = l.im;
        }
        for (int64_t c = j + 1; c <= ju; c++) {
          double *up = &AB(j, c);
          cplx u = {up[0], up[1]};
          if (u.re == 0.0 && u.im == 0.0) continue;
          for (int64_t i = 1; i <= km; i++) {
            double *lq = &AB(j + i, j), *tq = &AB(j + i, c);
            cplx t = csub((cplx){tq[0], tq[1]}, cmul((cplx){lq[0], lq[1]}, u));
            tq[0] = t.re;
            tq[1] = t.im;
          }
        }
      }
    }
  }
  int rc = OHM_OK;
  if (singular) {
    rc = fail(OHM_ERR_SOLVE, "Solve error: sparse LU factorization failed: singular matrix");
  } else {
    double *y = (double *)xmalloc(sizeof(double) * n * w);
    for (int64_t i = 0; i < n; i++)
      for (int t = 0; t < w; t++) y[iperm[i] * w + t] = b[i * w + t];
    /* L solve with row interchanges */
    for (int64_t j = 0; j < n; j++) {
      int64_t lm = kl < n - 1 - j ? kl : n - 1 - j;
      int64_t l = ipiv[j];
      if (l != j)
        for (int t = 0; t < w; t++) {
          double s = y[l * w + t];
          y[l * w + t] = y[j * w + t];
          y[j * w + t] = s;
        }
      if (!cplx_vals) {
        double yj = y[j];
        for (int64_t i = 1; i <= lm; i++) y[j + i] -= AB(j + i, j) * yj;
      } else {
        cplx yj = {y[2 * j], y[2 * j + 1]};
        for (int64_t i = 1; i <= lm; i++) {
          double *lq = &AB(j + i, j);
          cplx t = csub((cplx){y[2 * (j + i)], y[2 * (j + i) + 1]}, cmul((cplx){lq[0], lq[1]}, yj));
          y[2 * (j + i)] = t.re;
          y[2 * (j + i) + 1] = t.im;
        }
      }
    }
    /* U solve (upper bandwidth kv) */
    for (int64_t j = n - 1; j >= 0; j--) {
      int64_t i0 = j - kv > 0 ? j - kv : 0;
      if (!cplx_vals) {
        y[j] /= AB(j, j);
        double yj = y[j];
        for (int64_t i = i0; i < j; i++) y[i] -= AB(i, j) * yj;
      } else {
        double *d = &AB(j, j);
        cplx yj = cdiv((cplx){y[2 * j], y[2 * j + 1]}, (cplx){d[0], d[1]});
        y[2 * j] = yj.re;
        y[2 * j + 1] = yj.im;
        for (int64_t i = i0; i < j; i++) {
          double *u = &AB(i, j);
          cplx t = csub((cplx){y[2 * i], y[2 * i + 1]}, cmul((cplx){u[0], u[1]}, yj));
          y[2 * i] = t.re;
          y[2 * i + 1] = t.im;
        }
      }
    }
    for (int64_t i = 0; i < n; i++)
      for (int t = 0; t < w; t++) x[i * w + t] = y[iperm[i] * w + t];
    for (int64_t i = 0; i < n * w; i++)
      if (!isfinite(x[i])) {
        rc = fail(OHM_ERR_SOLVE, "Solve error: sparse LU produced NaN/Inf solution");
        break;
      }
    free(y);
  }
#undef AB
  free(ab);
  free(ipiv);
  free(iperm);
  return rc;
}

int oo_sparse_lu_solve(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                       const double *b, double *x, int cplx_vals) {
  if (n == 0) return OHM_OK;
  int64_t *perm = (int64_t *)xmalloc(sizeof(int64_t) * n);
  rcm_order(n, rowptr, col, perm);
  int rc = band_lu_solve(n, rowptr, col, val, b, x, cplx_vals, perm);
  free(perm);
  return rc;
}

/* ac::run over one circuit (src/analysis/ac.rs:56-70) with the sparse LU as
 * the linear solve: for every frequency f, A = G + j 2 pi f C
 * (form_complex_matrix, src/sparse.rs:149-178: G's and C's entries summed
 * position by position on the union of the two patterns, rows in ascending
 * column order) and x_f = A^-1 b_ac. The union pattern does not depend on f,
 * so its ordering is computed once. x: n_freqs * n complex, frequency-major. */
int oo_ac_sweep(int64_t n, const uint32_t *g_rp, const uint32_t *g_col, const double *g_val,
                const uint32_t *c_rp, const uint32_t *c_col, const double *c_val,
                const double *b_ac, int64_t n_freqs, const double *freqs, double *x) {
  if (n == 0 || n_freqs == 0) return OHM_OK;
  /* union pattern: merge the two sorted rows */
  int64_t cap = (int64_t)g_rp[n] + (c_val ? (int64_t)c_rp[n] : 0);
  uint32_t *rp = (uint32_t *)xmalloc(sizeof(uint32_t) * (n + 1));
  uint32_t *cl = (uint32_t *)xcalloc(cap > 0 ? cap : 1, sizeof(uint32_t));
  double *gv = (double *)xcalloc(cap > 0 ? cap : 1, sizeof(double));
  double *cv = (double *)xcalloc(cap > 0 ? cap : 1, sizeof(double));
  int64_t k = 0;
  rp[0] = 0;
  for (int64_t r = 0; r < n; r++) {
    uint32_t a = g_rp[r], ae = g_rp[r + 1];
    uint32_t c = c_val ? c_rp[r] : 0, ce = c_val ? c_rp[r + 1] : 0;
    while (a < ae || c < ce) {
      uint32_t ca = a < ae ? g_col[a] : 0xffffffffu, cc = c < ce ? c_col[c] : 0xffffffffu;
      uint32_t j = ca < cc ? ca : cc;
      cl[k] = j; /* a repeated column (never from the compiler) is summed by the LU's scatter */
      gv[k] = ca == j ? g_val[a++] : 0.0;
      cv[k] = cc == j ? c_val[c++] : 0.0;
      k++;
    }
    rp[r + 1] = (uint32_t)k;
  }
  int64_t *perm = (int64_t *)xmalloc(sizeof(int64_t) * n);
  rcm_order(n, rp, cl, perm);
  double *av = (double *)xmalloc(sizeof(double) * 2 * (k > 0 ? k : 1));
  int rc = OHM_OK;
  for (int64_t f = 0; f < n_freqs && rc == OHM_OK; f++) {
    const double om = 2.0 * 3.14159265358979323846 * freqs[f]; /* ac.rs:57 */
    for (int64_t q = 0; q < k; q++) {
      av[2 * q] = gv[q];
      av[2 * q + 1] = om * cv[q];
    }
    rc = band_lu_solve(n, rp, cl, av, b_ac, x + 2 * n * f, 1, perm);
  }
  free(av);
  free(perm);
  free(rp);
  free(cl);
  free(gv);
  free(cv);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* BiCGSTAB — src/solver/bicgstab.rs:56-189 (x0 = 0, right preconditioning). */
/* ------------------------------------------------------------------------ */
void oo_jacobi_inv_diag(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                        double *inv) {
  for (int64_t r = 0; r < n; r++) {
    inv[r] = 0.0;
    for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++)
      if ((int64_t)col[k] == r) {
        if (fabs(val[k]) > 1e-30) inv[r] = 1.0 / val[k];
        break;
      }
  }
}

static double dot(int64_t n, const double *a, const double *b) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}
static void axpy(int64_t n, double a, const double *x, double *y) {
  for (int64_t i = 0; i < n; i++) y[i] += a * x[i];
}

static void prec_apply(int64_t n, const oo_precond *p, const double *in, double *out, double *tmp) {
  if (!p || p->kind == OO_PREC_NONE) {
    memcpy(out, in, sizeof(double) * n);
  } else if (p->kind == OO_PREC_JACOBI) {
    for (int64_t i = 0; i < n; i++) out[i] = p->inv_diag[i] * in[i];
  } else {
    oo_spmv(n, p->ml_rowptr, p->ml_col, p->ml_val, in, tmp, 0);
    oo_spmv(n, p->mu_rowptr, p->mu_col, p->mu_val, tmp, out, 0);
  }
}

/* Dense shadow residual r_hat_i in [-1, 1): SplitMix64 finaliser of (i+1)*gamma. */
double oo_shadow_value(int64_t i) {
  uint64_t z = (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

int oo_bicgstab(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                const double *b, double *x, const oo_precond *p, int64_t max_iters, double tol,
                int shadow, int scaled, int64_t *iters) {
  *iters = 0;
  if (max_iters == 0) return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB max_iterations must be > 0");
  if (!(tol > 0.0) || !isfinite(tol))
    return fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB tolerance must be finite and > 0");
  double *buf = (double *)xcalloc(9 * n + 1, sizeof(double));
  double *r = buf, *rh = buf + n, *pp = buf + 2 * n, *v = buf + 3 * n, *s = buf + 4 * n,
         *t = buf + 5 * n, *ph = buf + 6 * n, *sh = buf + 7 * n, *tmp = buf + 8 * n;
  for (int64_t i = 0; i < n; i++) x[i] = 0.0;
  memcpy(r, b, sizeof(double) * n);
  if (shadow)
    for (int64_t i = 0; i < n; i++) rh[i] = oo_shadow_value(i);
  else
    memcpy(rh, b, sizeof(double) * n);
  double bnorm = sqrt(dot(n, r, r));
  int rc = OHM_OK;
  if (bnorm < 1e-30) goto done;
  double abs_tol = tol * bnorm, rho = 1.0, alpha = 1.0, omega = 1.0;
  double hh = dot(n, rh, rh);
  for (int64_t it = 0; it < max_iters; it++) {
    double rho_new = dot(n, rh, r);
    if (scaled ? fabs(rho_new) <= 1e-30 * (sqrt(hh) * sqrt(dot(n, r, r))) : fabs(rho_new) < 1e-30) {
      rc = fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB breakdown: rho ~ 0");
      goto done;
    }
    double beta = (rho_new / rho) * (alpha / omega);
    rho = rho_new;
    axpy(n, -omega, v, pp);
    for (int64_t i = 0; i < n; i++) pp[i] *= beta;
    axpy(n, 1.0, r, pp);
    prec_apply(n, p, pp, ph, tmp);
    oo_spmv(n, rowptr, col, val, ph, v, 0);
    double rv = dot(n, rh, v);
    if (scaled ? fabs(rv) <= 1e-30 * (sqrt(hh) * sqrt(dot(n, v, v))) : fabs(rv) < 1e-30) {
      rc = fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB breakdown: r_hat.v ~ 0");
      goto done;
    }
    alpha = rho / rv;
    memcpy(s, r, sizeof(double) * n);
    axpy(n, -alpha, v, s);
    double snorm = sqrt(dot(n, s, s));
    if (snorm < abs_tol) {
      axpy(n, alpha, ph, x);
      *iters = it + 1;
      goto done;
    }
    prec_apply(n, p, s, sh, tmp);
    oo_spmv(n, rowptr, col, val, sh, t, 0);
    double ts = dot(n, t, s), tt = dot(n, t, t);
    if (scaled ? tt <= 1e-30 * snorm * snorm : fabs(tt) < 1e-30) {
      rc = fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB breakdown: ||t|| ~ 0");
      goto done;
    }
    omega = ts / tt;
    axpy(n, alpha, ph, x);
    axpy(n, omega, sh, x);
    memcpy(r, s, sizeof(double) * n);
    axpy(n, -omega, t, r);
    double rnorm = sqrt(dot(n, r, r));
    if (isnan(rnorm) || isinf(rnorm)) {
      rc = fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB diverged: NaN/Inf in residual");
      goto done;
    }
    if (rnorm < abs_tol) {
      *iters = it + 1;
      goto done;
    }
    if (fabs(omega) < 1e-30) {
      rc = fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB breakdown: omega ~ 0");
      goto done;
    }
    *iters = it + 1;
  }
  rc = fail(OHM_ERR_SOLVE, "Solve error: BiCGSTAB did not converge within %lld iterations",
            (long long)max_iters);
done:
  free(buf);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Reference ISAI — src/solver/preconditioner.rs:64-428.                     */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t n, cap;
  int64_t *c;
  double *v;
} rowlist;

static void rl_reserve(rowlist *r, int64_t need) {
  if (need <= r->cap) return;
  int64_t cap = r->cap ? r->cap : 8;
  while (cap < need) cap *= 2;
  r->c = (int64_t *)realloc(r->c, sizeof(int64_t) * cap);
  r->v = (double *)realloc(r->v, sizeof(double) * cap);
  if (!r->c || !r->v) abort();
  r->cap = cap;
}
static int64_t rl_find(const rowlist *r, int64_t c) {
  for (int64_t k = 0; k < r->n; k++)
    if (r->c[k] == c) return k;
  return -1;
}
static double rl_val(const rowlist *r, int64_t c) {
  int64_t k = rl_find(r, c);
  return k < 0 ? 0.0 : r->v[k];
}
static void rl_insert_sorted(rowlist *r, int64_t c, double v) {
  int64_t pos = 0;
  while (pos < r->n && r->c[pos] < c) pos++; /* partition_point */
  rl_reserve(r, r->n + 1);
  memmove(r->c + pos + 1, r->c + pos, sizeof(int64_t) * (r->n - pos));
  memmove(r->v + pos + 1, r->v + pos, sizeof(double) * (r->n - pos));
  r->c[pos] = c;
  r->v[pos] = v;
  r->n++;
}

/* CSR built through from_triplets from a triplet list (keeps explicit zeros). */
static void csr_from_trip(int64_t n, tripvec *t, int64_t *nnz, uint32_t **rp, uint32_t **cl,
                          double **vv) {
  int64_t *rp64 = (int64_t *)xmalloc(sizeof(int64_t) * (n + 1));
  int64_t *c64 = (int64_t *)xmalloc(sizeof(int64_t) * (t->n + 1));
  double *v = (double *)xmalloc(sizeof(double) * (t->n + 1));
  int64_t z = oo_from_triplets(n, t->n, t->r, t->c, t->v, 0, rp64, c64, v);
  *rp = (uint32_t *)xmalloc(sizeof(uint32_t) * (n + 1));
  *cl = (uint32_t *)xmalloc(sizeof(uint32_t) * (z + 1));
  for (int64_t i = 0; i <= n; i++) (*rp)[i] = (uint32_t)rp64[i];
  for (int64_t k = 0; k < z; k++) (*cl)[k] = (uint32_t)c64[k];
  *vv = v;
  *nnz = z;
  free(rp64);
  free(c64);
}

/* ilu0 (:64-153): column-major LU with partial pivoting and fill. */
static void ref_lu(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                   rowlist **rows_out, int64_t **perm_out) {
  rowlist *rows = (rowlist *)xcalloc(n + 1, sizeof(rowlist));
  int64_t *perm = (int64_t *)xmalloc(sizeof(int64_t) * (n + 1));
  for (int64_t i = 0; i < n; i++) {
    perm[i] = i;
    rl_reserve(&rows[i], rowptr[i + 1] - rowptr[i] + 1);
    for (uint32_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
      rows[i].c[rows[i].n] = col[k];
      rows[i].v[rows[i].n] = val[k];
      rows[i].n++;
    }
  }
  rowlist pu = {0};
  for (int64_t j = 0; j < n; j++) {
    int64_t best_row = j;
    double best_val = fabs(rl_val(&rows[j], j));
    for (int64_t i = j + 1; i < n; i++) {
      double v = fabs(rl_val(&rows[i], j));
      if (v > best_val) {
        best_val = v;
        best_row = i;
      }
    }
    if (best_row != j) {
      rowlist t = rows[j];
      rows[j] = rows[best_row];
      rows[best_row] = t;
      int64_t tp = perm[j];
      perm[j] = perm[best_row];
      perm[best_row] = tp;
    }
    double pivot = rl_val(&rows[j], j);
    if (fabs(pivot) < 1e-30) continue;
    pu.n = 0;
    for (int64_t k = 0; k < rows[j].n; k++)
      if (rows[j].c[k] > j) {
        rl_reserve(&pu, pu.n + 1);
        pu.c[pu.n] = rows[j].c[k];
        pu.v[pu.n] = rows[j].v[k];
        pu.n++;
      }
    for (int64_t i = j + 1; i < n; i++) {
      double a_ij = rl_val(&rows[i], j);
      if (fabs(a_ij) < 1e-30) continue;
      double l_ij = a_ij / pivot;
      int64_t pos = rl_find(&rows[i], j); /* set_or_insert (:156-163) */
      if (pos >= 0)
        rows[i].v[pos] = l_ij;
      else
        rl_insert_sorted(&rows[i], j, l_ij);
      for (int64_t q = 0; q < pu.n; q++) {
        double fillv = l_ij * pu.v[q];
        int64_t p2 = rl_find(&rows[i], pu.c[q]);
        if (p2 >= 0)
          rows[i].v[p2] = rows[i].v[p2] - fillv;
        else
          rl_insert_sorted(&rows[i], pu.c[q], 0.0 - fillv);
      }
    }
  }
  free(pu.c);
  free(pu.v);
  *rows_out = rows;
  *perm_out = perm;
}

static void rows_to_lu(int64_t n, rowlist *rows, tripvec *lt, tripvec *ut) {
  for (int64_t i = 0; i < n; i++) {
    for (int64_t k = 0; k < rows[i].n; k++)
      if (rows[i].c[k] < i) tv_push(lt, i, rows[i].c[k], rows[i].v[k]);
    tv_push(lt, i, i, 1.0);
    for (int64_t k = 0; k < rows[i].n; k++)
      if (rows[i].c[k] >= i) tv_push(ut, i, rows[i].c[k], rows[i].v[k]);
  }
}

int oo_lu_reference(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                    int64_t *l_nnz, uint32_t **l_rowptr, uint32_t **l_col, double **l_val,
                    int64_t *u_nnz, uint32_t **u_rowptr, uint32_t **u_col, double **u_val,
                    int64_t **perm) {
  rowlist *rows;
  ref_lu(n, rowptr, col, val, &rows, perm);
  tripvec lt = {0}, ut = {0};
  rows_to_lu(n, rows, &lt, &ut);
  csr_from_trip(n, &lt, l_nnz, l_rowptr, l_col, l_val);
  csr_from_trip(n, &ut, u_nnz, u_rowptr, u_col, u_val);
  for (int64_t i = 0; i < n; i++) {
    free(rows[i].c);
    free(rows[i].v);
  }
  free(rows);
  tv_free(&lt);
  tv_free(&ut);
  return OHM_OK;
}

/* Symbolic pattern of A*B, sorted + dedup per row (:185-201). Returns CSR pattern. */
static void sym_mult(int64_t n, const uint32_t *arp, const uint32_t *acl, const uint32_t *brp,
                     const uint32_t *bcl, uint32_t **rp_out, uint32_t **cl_out) {
  uint32_t *rp = (uint32_t *)xmalloc(sizeof(uint32_t) * (n + 1));
  int64_t cap = 1024, nz = 0;
  uint32_t *cl = (uint32_t *)xmalloc(sizeof(uint32_t) * cap);
  int64_t *mark = (int64_t *)xmalloc(sizeof(int64_t) * (n + 1));
  for (int64_t i = 0; i < n; i++) mark[i] = -1;
  rp[0] = 0;
  for (int64_t i = 0; i < n; i++) {
    int64_t start = nz;
    for (uint32_t a = arp[i]; a < arp[i + 1]; a++) {
      uint32_t k = acl[a];
      for (uint32_t b = brp[k]; b < brp[k + 1]; b++) {
        uint32_t j = bcl[b];
        if (mark[j] != i) {
          mark[j] = i;
          if (nz == cap) {
            cap *= 2;
            cl = (uint32_t *)realloc(cl, sizeof(uint32_t) * cap);
          }
          cl[nz++] = j;
        }
      }
    }
    for (int64_t x = start + 1; x < nz; x++) {
      uint32_t v = cl[x];
      int64_t y = x - 1;
      while (y >= start && cl[y] > v) {
        cl[y + 1] = cl[y];
        y--;
      }
      cl[y + 1] = v;
    }
    rp[i + 1] = (uint32_t)nz;
  }
  free(mark);
  *rp_out = rp;
  *cl_out = cl;
}

/* Column pattern (transpose of a row pattern), rows sorted ascending (:294-306). */
static void transpose_pattern(int64_t n, const uint32_t *rp, const uint32_t *cl, uint32_t **crp,
                              uint32_t **crow) {
  uint32_t *c = (uint32_t *)xcalloc(n + 2, sizeof(uint32_t));
  for (int64_t k = 0; k < (int64_t)rp[n]; k++) c[cl[k] + 1]++;
  for (int64_t i = 0; i < n; i++) c[i + 1] += c[i];
  uint32_t *pos = (uint32_t *)xmalloc(sizeof(uint32_t) * (n + 1));
  for (int64_t i = 0; i < n; i++) pos[i] = c[i];
  uint32_t *rows = (uint32_t *)xmalloc(sizeof(uint32_t) * (rp[n] + 1));
  for (int64_t r = 0; r < n; r++)
    for (uint32_t k = rp[r]; k < rp[r + 1]; k++) rows[pos[cl[k]]++] = (uint32_t)r;
  free(pos);
  *crp = c;
  *crow = rows;
}

/* dense_solve (:371-428): GE with partial pivoting; zero pivots leave zeros. */
static void dense_solve_small(int64_t n, double *a, double *b, double *x) {
  for (int64_t cidx = 0; cidx < n; cidx++) {
    int64_t best_row = cidx;
    double best_val = fabs(a[cidx * n + cidx]);
    for (int64_t r = cidx + 1; r < n; r++) {
      double v = fabs(a[r * n + cidx]);
      if (v > best_val) {
        best_val = v;
        best_row = r;
      }
    }
    if (best_val < 1e-30) continue;
    if (best_row != cidx) {
      for (int64_t j = 0; j < n; j++) {
        double t = a[cidx * n + j];
        a[cidx * n + j] = a[best_row * n + j];
        a[best_row * n + j] = t;
      }
      double t = b[cidx];
      b[cidx] = b[best_row];
      b[best_row] = t;
    }
    double pivot = a[cidx * n + cidx];
    for (int64_t r = cidx + 1; r < n; r++) {
      double f = a[r * n + cidx] / pivot;
      a[r * n + cidx] = 0.0;
      for (int64_t j = cidx + 1; j < n; j++) a[r * n + j] = a[r * n + j] - f * a[cidx * n + j];
      b[r] = b[r] - f * b[cidx];
    }
  }
  for (int64_t cidx = n - 1; cidx >= 0; cidx--) {
    double d = a[cidx * n + cidx];
    if (fabs(d) < 1e-30) {
      x[cidx] = 0.0;
      continue;
    }
    double s = b[cidx];
    for (int64_t j = cidx + 1; j < n; j++) s = s - a[cidx * n + j] * x[j];
    x[cidx] = s / d;
  }
}

/* solve_column_isai (:316-367) for every column, assembled via columns_to_csr (:431-444). */
static void isai_columns(int64_t n, const uint32_t *frp, const uint32_t *fcl, const double *fv,
                         const uint32_t *crp, const uint32_t *crow, tripvec *out) {
  int64_t *g2l = (int64_t *)xmalloc(sizeof(int64_t) * (n + 1));
  for (int64_t i = 0; i < n; i++) g2l[i] = -1;
  for (int64_t k = 0; k < n; k++) {
    int64_t nz = crp[k + 1] - crp[k];
    if (nz == 0) continue;
    const uint32_t *J = crow + crp[k];
    for (int64_t l = 0; l < nz; l++) g2l[J[l]] = l;
    double *sub = (double *)xcalloc(nz * nz, sizeof(double));
    double *rhs = (double *)xcalloc(nz, sizeof(double));
    double *sol = (double *)xcalloc(nz, sizeof(double));
    for (int64_t li = 0; li < nz; li++) {
      uint32_t r = J[li];
      for (uint32_t q = frp[r]; q < frp[r + 1]; q++)
        if (g2l[fcl[q]] >= 0) sub[li * nz + g2l[fcl[q]]] = fv[q];
    }
    if (g2l[k] >= 0) rhs[g2l[k]] = 1.0;
    dense_solve_small(nz, sub, rhs, sol);
    for (int64_t l = 0; l < nz; l++) {
      tv_push(out, J[l], k, sol[l]);
      g2l[J[l]] = -1;
    }
    free(sub);
    free(rhs);
    free(sol);
  }
  free(g2l);
}

int oo_isai_reference(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                      int level, int64_t *ml_nnz, uint32_t **ml_rowptr, uint32_t **ml_col,
                      double **ml_val, int64_t *mu_nnz, uint32_t **mu_rowptr, uint32_t **mu_col,
                      double **mu_val, int64_t **perm_out) {
  if (level != 0 && level != 1) return fail(OHM_ERR_INVALID, "ISAI level must be 0 or 1");
  int64_t lnz, unz;
  uint32_t *lrp, *lcl, *urp, *ucl;
  double *lv, *uv;
  int64_t *perm;
  oo_lu_reference(n, rowptr, col, val, &lnz, &lrp, &lcl, &lv, &unz, &urp, &ucl, &uv, &perm);
  uint32_t *mlp_rp, *mlp_cl, *mup_rp, *mup_cl;
  if (level == 0) {
    mlp_rp = (uint32_t *)xmalloc(sizeof(uint32_t) * (n + 1));
    memcpy(mlp_rp, lrp, sizeof(uint32_t) * (n + 1));
    mlp_cl = (uint32_t *)xmalloc(sizeof(uint32_t) * (lnz + 1));
    memcpy(mlp_cl, lcl, sizeof(uint32_t) * lnz);
    mup_rp = (uint32_t *)xmalloc(sizeof(uint32_t) * (n + 1));
    memcpy(mup_rp, urp, sizeof(uint32_t) * (n + 1));
    mup_cl = (uint32_t *)xmalloc(sizeof(uint32_t) * (unz + 1));
    memcpy(mup_cl, ucl, sizeof(uint32_t) * unz);
  } else {
    sym_mult(n, lrp, lcl, lrp, lcl, &mlp_rp, &mlp_cl);
    sym_mult(n, urp, ucl, urp, ucl, &mup_rp, &mup_cl);
  }
  uint32_t *crp, *crow;
  tripvec mlt = {0}, mut = {0};
  transpose_pattern(n, mlp_rp, mlp_cl, &crp, &crow);
  isai_columns(n, lrp, lcl, lv, crp, crow, &mlt);
  free(crp);
  free(crow);
  /* fold the row permutation into M_L's columns: new_col = perm[old_col] (:280-291) */
  for (int64_t t = 0; t < mlt.n; t++) mlt.c[t] = perm[mlt.c[t]];
  transpose_pattern(n, mup_rp, mup_cl, &crp, &crow);
  isai_columns(n, urp, ucl, uv, crp, crow, &mut);
  free(crp);
  free(crow);
  csr_from_trip(n, &mlt, ml_nnz, ml_rowptr, ml_col, ml_val);
  csr_from_trip(n, &mut, mu_nnz, mu_rowptr, mu_col, mu_val);
  *perm_out = perm;
  free(lrp);
  free(lcl);
  free(lv);
  free(urp);
  free(ucl);
  free(uv);
  free(mlp_rp);
  free(mlp_cl);
  free(mup_rp);
  free(mup_cl);
  tv_free(&mlt);
  tv_free(&mut);
  return OHM_OK;
}

/* ------------------------------------------------------------------------ */
/* Backend preconditioner computed exactly: ILU(0) (Saad, "Iterative Methods */
/* for Sparse Linear Systems", Alg. 10.4, IKJ form) on pattern(A) + diagonal */
/* with no pivoting; a multiplier whose pivot is below 1e-30 is set to zero. */
/* ISAI columns solve L[J,J] m = e_k / U[J,J] m = e_k by substitution.       */
/* ------------------------------------------------------------------------ */
int oo_ilu0_isai(int64_t n, const uint32_t *rowptr, const uint32_t *col, const double *val,
                 int level, int64_t *l_nnz, uint32_t **l_rowptr, uint32_t **l_col, double **l_val,
                 int64_t *u_nnz, uint32_t **u_rowptr, uint32_t **u_col, double **u_val,
                 int64_t *ml_nnz, uint32_t **ml_rowptr, uint32_t **ml_col, double **ml_val,
                 int64_t *mu_nnz, uint32_t **mu_rowptr, uint32_t **mu_col, double **mu_val) {
  if (level < 0 || level > 1) return fail(OHM_ERR_INVALID, "ISAI level must be 0 or 1");
  /* factor pattern = pattern(A) U diag */
  uint32_t *frp = (uint32_t *)xmalloc(sizeof(uint32_t) * (n + 1));
  uint32_t *fcl = (uint32_t *)xmalloc(sizeof(uint32_t) * (rowptr[n] + n + 1));
  double *fv = (double *)xmalloc(sizeof(double) * (rowptr[n] + n + 1));
  int64_t nz = 0;
  frp[0] = 0;
  for (int64_t r = 0; r < n; r++) {
    int have = 0;
    for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
      if (!have && (int64_t)col[k] > r) {
        fcl[nz] = (uint32_t)r;
        fv[nz++] = 0.0;
        have = 1;
      }
      if ((int64_t)col[k] == r) have = 1;
      fcl[nz] = col[k];
      fv[nz++] = val[k];
    }
    if (!have) {
      fcl[nz] = (uint32_t)r;
      fv[nz++] = 0.0;
    }
    frp[r + 1] = (uint32_t)nz;
  }
  int64_t *diag = (int64_t *)xmalloc(sizeof(int64_t) * (n + 1));
  int64_t *pos = (int64_t *)xmalloc(sizeof(int64_t) * (n + 1));
  for (int64_t i = 0; i < n; i++) pos[i] = -1;
  for (int64_t r = 0; r < n; r++)
    for (uint32_t k = frp[r]; k < frp[r + 1]; k++)
      if ((int64_t)fcl[k] == r) diag[r] = k;
  for (int64_t i = 0; i < n; i++) {
    for (uint32_t k = frp[i]; k < frp[i + 1]; k++) pos[fcl[k]] = k;
    for (uint32_t k = frp[i]; k < frp[i + 1] && (int64_t)fcl[k] < i; k++) {
      int64_t kk = fcl[k];
      double piv = fv[diag[kk]];
      if (fabs(piv) < 1e-30) {
        fv[k] = 0.0;
        continue;
      }
      double l = fv[k] / piv;
      fv[k] = l;
      for (uint32_t q = diag[kk] + 1; q < frp[kk + 1]; q++)
        if (pos[fcl[q]] >= 0) fv[pos[fcl[q]]] -= l * fv[q];
    }
    for (uint32_t k = frp[i]; k < frp[i + 1]; k++) pos[fcl[k]] = -1;
  }
  tripvec lt = {0}, ut = {0};
  for (int64_t r = 0; r < n; r++) {
    for (uint32_t k = frp[r]; k < frp[r + 1]; k++) {
      if ((int64_t)fcl[k] < r) tv_push(&lt, r, fcl[k], fv[k]);
      if ((int64_t)fcl[k] == r) tv_push(&lt, r, r, 1.0);
      if ((int64_t)fcl[k] >= r) tv_push(&ut, r, fcl[k], fv[k]);
    }
  }
  csr_from_trip(n, &lt, l_nnz, l_rowptr, l_col, l_val);
  csr_from_trip(n, &ut, u_nnz, u_rowptr, u_col, u_val);
  uint32_t *mlp_rp, *mlp_cl, *mup_rp, *mup_cl;
  if (level == 1) {
    sym_mult(n, *l_rowptr, *l_col, *l_rowptr, *l_col, &mlp_rp, &mlp_cl);
    sym_mult(n, *u_rowptr, *u_col, *u_rowptr, *u_col, &mup_rp, &mup_cl);
  } else {
    mlp_rp = (uint32_t *)xmalloc(sizeof(uint32_t) * (n + 1));
    memcpy(mlp_rp, *l_rowptr, sizeof(uint32_t) * (n + 1));
    mlp_cl = (uint32_t *)xmalloc(sizeof(uint32_t) * (*l_nnz + 1));
    memcpy(mlp_cl, *l_col, sizeof(uint32_t) * *l_nnz);
    mup_rp = (uint32_t *)xmalloc(sizeof(uint32_t) * (n + 1));
    memcpy(mup_rp, *u_rowptr, sizeof(uint32_t) * (n + 1));
    mup_cl = (uint32_t *)xmalloc(sizeof(uint32_t) * (*u_nnz + 1));
    memcpy(mup_cl, *u_col, sizeof(uint32_t) * *u_nnz);
  }
  /* M_L column k: J ascending with J[0] = k; unit-diagonal forward substitution. */
  tripvec mlt = {0}, mut = {0};
  uint32_t *crp, *crow;
  int64_t *g2l = (int64_t *)xmalloc(sizeof(int64_t) * (n + 1));
  for (int64_t i = 0; i < n; i++) g2l[i] = -1;
  transpose_pattern(n, mlp_rp, mlp_cl, &crp, &crow);
  double *m = (double *)xmalloc(sizeof(double) * (n + 1));
  const uint32_t *Lrp = *l_rowptr, *Lcl = *l_col;
  const double *Lv = *l_val;
  for (int64_t k = 0; k < n; k++) {
    const uint32_t *J = crow + crp[k];
    int64_t nzc = crp[k + 1] - crp[k];
    for (int64_t t = 0; t < nzc; t++) g2l[J[t]] = t;
    for (int64_t t = 0; t < nzc; t++) {
      int64_t r = J[t];
      double s = (r == k) ? 1.0 : 0.0;
      double d = 1.0;
      for (uint32_t q = Lrp[r]; q < Lrp[r + 1]; q++) {
        int64_t c = Lcl[q];
        if (c == r) {
          d = Lv[q];
        } else if (g2l[c] >= 0 && g2l[c] < t) {
          s -= Lv[q] * m[g2l[c]];
        }
      }
      m[t] = fabs(d) < 1e-30 ? 0.0 : s / d;
    }
    for (int64_t t = 0; t < nzc; t++) {
      tv_push(&mlt, J[t], k, m[t]);
      g2l[J[t]] = -1;
    }
  }
  free(crp);
  free(crow);
  /* M_U column k: J ascending with J[last] = k; back substitution. */
  transpose_pattern(n, mup_rp, mup_cl, &crp, &crow);
  const uint32_t *Urp = *u_rowptr, *Ucl = *u_col;
  const double *Uv = *u_val;
  for (int64_t k = 0; k < n; k++) {
    const uint32_t *J = crow + crp[k];
    int64_t nzc = crp[k + 1] - crp[k];
    for (int64_t t = 0; t < nzc; t++) g2l[J[t]] = t;
    for (int64_t t = nzc - 1; t >= 0; t--) {
      int64_t r = J[t];
      double s = (r == k) ? 1.0 : 0.0;
      double d = 0.0;
      for (uint32_t q = Urp[r]; q < Urp[r + 1]; q++) {
        int64_t c = Ucl[q];
        if (c == r) {
          d = Uv[q];
        } else if (g2l[c] >= 0 && g2l[c] > t) {
          s -= Uv[q] * m[g2l[c]];
        }
      }
      m[t] = fabs(d) < 1e-30 ? 0.0 : s / d;
    }
    for (int64_t t = 0; t < nzc; t++) {
      tv_push(&mut, J[t], k, m[t]);
      g2l[J[t]] = -1;
    }
  }
  free(crp);
  free(crow);
  csr_from_trip(n, &mlt, ml_nnz, ml_rowptr, ml_col, ml_val);
  csr_from_trip(n, &mut, mu_nnz, mu_rowptr, mu_col, mu_val);
  free(m);
  free(g2l);
  free(frp);
  free(fcl);
  free(fv);
  free(diag);
  free(pos);
  free(mlp_rp);
  free(mlp_cl);
  free(mup_rp);
  free(mup_cl);
  tv_free(&lt);
  tv_free(&ut);
  tv_free(&mlt);
  tv_free(&mut);
  return OHM_OK;
}

/* ------------------------------------------------------------------------ */
/* Device models — FP64 restatements of src/solver/ds_nonlinear_shaders.rs.  */
/* ------------------------------------------------------------------------ */
static inline double xv(const double *x, uint32_t idx) { return idx == OHM_SENTINEL ? 0.0 : x[idx]; }
static inline double clamp_exp(double a) { /* ds_exp clamps to [-80, 80] (ds_shaders.rs:96-108) */
  if (a > 80.0) a = 80.0;
  if (a < -80.0) a = -80.0;
  return exp(a);
}

/* diode_eval_ds (:41-76) */
void oo_diode_eval(const ohm_diode_desc *d, const double *x, double *i_out, double *g_out) {
  double is = d->is_val, nvt = d->n_vt;
  double vd = xv(x, d->anode_idx) - xv(x, d->cathode_idx);
  double e = clamp_exp(vd / nvt);
  *i_out = is * (e - 1.0);
  *g_out = is * e / nvt;
}

/* bjt_eval_ds (:281-367) */
void oo_bjt_eval(const ohm_bjt_desc *d, const double *x, double *out) {
  double pol = d->polarity, is = d->is_val, bf = d->bf, br = d->br, nf = d->nf_vt, nr = d->nr_vt;
  double vb = xv(x, d->base_idx), ve = xv(x, d->emitter_idx), vc = xv(x, d->collector_idx);
  double vbe = pol * (vb - ve), vbc = pol * (vb - vc);
  double ef = clamp_exp(vbe / nf), er = clamp_exp(vbc / nr);
  double i_f = is * (ef - 1.0), i_r = is * (er - 1.0);
  double bf_frac = bf / (bf + 1.0);
  double ic = bf_frac * i_f - i_r / (br + 1.0);
  double ib = i_f / (bf + 1.0) + i_r / (br + 1.0);
  double gf = is * ef / nf, gr = is * er / nr;
  out[0] = ic * pol;
  out[1] = ib * pol;
  out[2] = bf_frac * gf;        /* dIc/dVbe */
  out[3] = -(gr / (br + 1.0));  /* dIc/dVbc */
  out[4] = gf / (bf + 1.0);     /* dIb/dVbe */
  out[5] = gr / (br + 1.0);     /* dIb/dVbc */
}

/* mosfet_eval_ds (:747-849); CPU twin src/solver/gpu_shaders.rs:1452-1493. */
void oo_mosfet_core(double vgs, double vds, double vto, double kp, double lambda, double polarity,
                    double *id, double *gm, double *gds) {
  const double VT = 0.02585;
  vgs *= polarity;
  vds *= polarity;
  double v_ov = vgs - vto, arg = v_ov / VT, v_ov_eff, sig;
  if (arg > 20.0) {
    v_ov_eff = v_ov;
    sig = 1.0;
  } else if (arg < -20.0) {
    double e = exp(arg);
    v_ov_eff = VT * e;
    sig = e;
  } else {
    v_ov_eff = VT * log1p(exp(arg));
    sig = 1.0 / (1.0 + exp(-arg));
  }
  double lam = 1.0 + lambda * vds;
  double i, g_m, g_ds;
  if (vds < v_ov_eff) {
    double inner = v_ov_eff * vds - vds * vds * 0.5;
    i = kp * inner * lam;
    g_m = kp * vds * lam * sig;
    g_ds = kp * (v_ov_eff - vds) * lam + kp * inner * lambda;
  } else {
    double sat = 0.5 * kp * v_ov_eff * v_ov_eff;
    i = sat * lam;
    g_m = kp * v_ov_eff * lam * sig;
    g_ds = sat * lambda;
  }
  *id = polarity * i;
  *gm = g_m;
  *gds = g_ds;
}

void oo_mosfet_eval(const ohm_mosfet_desc *d, const double *x, double *out) {
  double vg = xv(x, d->gate_idx), vs = xv(x, d->source_idx), vdd = xv(x, d->drain_idx);
  oo_mosfet_core(vg - vs, vdd - vs, d->vto, d->kp, d->lambda, d->polarity, &out[0], &out[1],
                 &out[2]);
}

/* pnjlim exactly as written in the reference (:615-629, :1173-1187). */
double oo_pnjlim(double vnew, double vold, double vt, double vcrit) {
  if (vnew > vcrit && fabs(vnew - vold) > vt + vt) {
    if (vold > 0.0) {
      double arg = 1.0 + (vnew - vold) / vt;
      if (arg > 0.0) return vold + vt * (2.0 + log(arg - 2.0));
      return vcrit;
    }
    return vt * log(vnew / vt);
  }
  return vnew;
}

static inline void addslot(double *v, uint32_t s, double a) {
  if (s != OHM_SENTINEL) v[s] += a;
}

/* Assembly: base copy + stamps. Diode matrix stamp uses the physical signs
 * for the compiler's slot order [aa, cc, ac, ca] (src/compiler.rs:59-63);
 * RHS Norton stamps follow :212-248 (diode), :477-573 (BJT), :936-996 (MOSFET). */
void oo_newton_assemble(int64_t n, int64_t nnz, const double *base_val, const double *base_b,
                        int64_t n_diodes, const ohm_diode_desc *diodes, int64_t n_bjts,
                        const ohm_bjt_desc *bjts, int64_t n_mosfets,
                        const ohm_mosfet_desc *mosfets, const double *x, double *val, double *b) {
  memcpy(val, base_val, sizeof(double) * nnz);
  memcpy(b, base_b, sizeof(double) * n);
  for (int64_t k = 0; k < n_diodes; k++) {
    const ohm_diode_desc *d = &diodes[k];
    double i, g;
    oo_diode_eval(d, x, &i, &g);
    addslot(val, d->g_row_col[0], g);
    addslot(val, d->g_row_col[1], g);
    addslot(val, d->g_row_col[2], -g);
    addslot(val, d->g_row_col[3], -g);
    double vd = xv(x, d->anode_idx) - xv(x, d->cathode_idx);
    double ieq = i - g * vd;
    addslot(b, d->b_idx[0], -ieq);
    addslot(b, d->b_idx[1], ieq);
  }
  for (int64_t k = 0; k < n_bjts; k++) {
    const ohm_bjt_desc *d = &bjts[k];
    double o[6];
    oo_bjt_eval(d, x, o);
    double ic = o[0], ib = o[1], a = o[2], bb = o[3], c = o[4], dd = o[5];
    double G[9] = {-bb, a + bb, -a, -dd, c + dd, -c, bb + dd, -(a + bb + c + dd), a + c};
    for (int j = 0; j < 9; j++) addslot(val, d->g_row_col[j], G[j]);
    double vc = xv(x, d->collector_idx), vb = xv(x, d->base_idx), ve = xv(x, d->emitter_idx);
    double ie = -(ic + ib);
    double ilc = G[0] * vc + G[1] * vb + G[2] * ve;
    double ilb = G[3] * vc + G[4] * vb + G[5] * ve;
    double ile = G[6] * vc + G[7] * vb + G[8] * ve;
    addslot(b, d->b_idx[0], -(ic - ilc));
    addslot(b, d->b_idx[1], -(ib - ilb));
    addslot(b, d->b_idx[2], -(ie - ile));
  }
  for (int64_t k = 0; k < n_mosfets; k++) {
    const ohm_mosfet_desc *d = &mosfets[k];
    double o[3];
    oo_mosfet_eval(d, x, o);
    double id = o[0], gm = o[1], gds = o[2];
    double S[6] = {gds, gm, -(gm + gds), -gds, -gm, gm + gds};
    for (int j = 0; j < 6; j++) addslot(val, d->g_row_col[j], S[j]);
    double vd = xv(x, d->drain_idx), vg = xv(x, d->gate_idx), vs = xv(x, d->source_idx);
    double ild = gds * vd + gm * vg - (gm + gds) * vs;
    double ils = -gds * vd - gm * vg + (gm + gds) * vs;
    addslot(b, d->b_idx[0], -(id - ild));
    addslot(b, d->b_idx[1], -(-id - ils));
  }
}

/* Apply x_new[t] = x_old[t] + s * (x_new[t] - x_old[t]) for one terminal. */
static inline void scale_node(const double *xo, double *xn, uint32_t t, double s) {
  if (t == OHM_SENTINEL) return;
  xn[t] = xo[t] + (xn[t] - xo[t]) * s;
}

/* Voltage limiting in the reference's order: BJT (:632-714), MOSFET two-pass
 * per-node minimum scale (:1052-1130), diode (:1190-1234). Devices are
 * applied in index order. */
void oo_newton_limit(int64_t n, int64_t n_diodes, const ohm_diode_desc *diodes, int64_t n_bjts,
                     const ohm_bjt_desc *bjts, int64_t n_mosfets, const ohm_mosfet_desc *mosfets,
                     const double *xo, double *xn) {
  const double SQRT2 = 1.4142135623730951;
  for (int64_t k = 0; k < n_bjts; k++) {
    const ohm_bjt_desc *d = &bjts[k];
    double pol = d->polarity, is = d->is_val, nf = d->nf_vt, nr = d->nr_vt;
    double vcf = nf * log(nf / (SQRT2 * is)), vcr = nr * log(nr / (SQRT2 * is));
    uint32_t B = d->base_idx, E = d->emitter_idx, C = d->collector_idx;
    double vo = pol * (xv(xo, B) - xv(xo, E)), vnw = pol * (xv(xn, B) - xv(xn, E));
    double vl = oo_pnjlim(vnw, vo, nf, vcf);
    if (vl != vnw) {
      double delta = vnw - vo;
      if (fabs(delta) > 1e-30) {
        double s = (vl - vo) / delta;
        if (!d->node_fixed[1]) scale_node(xo, xn, B, s);
        if (!d->node_fixed[2]) scale_node(xo, xn, E, s);
      }
    }
    vo = pol * (xv(xo, B) - xv(xo, C));
    vnw = pol * (xv(xn, B) - xv(xn, C));
    vl = oo_pnjlim(vnw, vo, nr, vcr);
    if (vl != vnw) {
      double delta = vnw - vo;
      if (fabs(delta) > 1e-30) {
        double s = (vl - vo) / delta;
        if (!d->node_fixed[1]) scale_node(xo, xn, B, s);
        if (!d->node_fixed[0]) scale_node(xo, xn, C, s);
      }
    }
  }
  if (n_mosfets > 0) {
    const uint32_t ONE = 1000000u; /* MVL_SCALE_ONE */
    const double STEP = 0.5;       /* MVL_STEP_LIMIT */
    uint32_t *sc = (uint32_t *)xmalloc(sizeof(uint32_t) * (n + 1));
    for (int64_t i = 0; i < n; i++) sc[i] = ONE;
    for (int64_t k = 0; k < n_mosfets; k++) {
      const ohm_mosfet_desc *d = &mosfets[k];
      uint32_t D = d->drain_idx, G = d->gate_idx, S = d->source_idx, f = d->node_fixed_packed;
      double dgs = (xv(xn, G) - xv(xn, S)) - (xv(xo, G) - xv(xo, S));
      if (fabs(dgs) > STEP) {
        double e = floor(STEP / fabs(dgs) * (double)ONE);
        uint32_t enc = e < 1.0 ? 1u : (uint32_t)e;
        if (G != OHM_SENTINEL && !(f & 2u) && enc < sc[G]) sc[G] = enc;
        if (S != OHM_SENTINEL && !(f & 4u) && enc < sc[S]) sc[S] = enc;
      }
      double dds = (xv(xn, D) - xv(xn, S)) - (xv(xo, D) - xv(xo, S));
      if (fabs(dds) > STEP) {
        double e = floor(STEP / fabs(dds) * (double)ONE);
        uint32_t enc = e < 1.0 ? 1u : (uint32_t)e;
        if (D != OHM_SENTINEL && !(f & 1u) && enc < sc[D]) sc[D] = enc;
        if (S != OHM_SENTINEL && !(f & 4u) && enc < sc[S]) sc[S] = enc;
      }
    }
    for (int64_t i = 0; i < n; i++)
      if (sc[i] < ONE) xn[i] = xo[i] + (xn[i] - xo[i]) * ((double)sc[i] / (double)ONE);
    free(sc);
  }
  for (int64_t k = 0; k < n_diodes; k++) {
    const ohm_diode_desc *d = &diodes[k];
    double nvt = d->n_vt, is = d->is_val;
    double vcrit = nvt * log(nvt / (SQRT2 * is));
    double vo = xv(xo, d->anode_idx) - xv(xo, d->cathode_idx);
    double vnw = xv(xn, d->anode_idx) - xv(xn, d->cathode_idx);
    double vl = oo_pnjlim(vnw, vo, nvt, vcrit);
    if (vl != vnw) {
      double delta = vnw - vo;
      if (fabs(delta) > 1e-30) {
        double s = (vl - vo) / delta;
        scale_node(xo, xn, d->anode_idx, s);
        scale_node(xo, xn, d->cathode_idx, s);
      }
    }
  }
}

/* newton_solve — src/solver/newton.rs:153-488, linear sub-problem by LU. */
int oo_newton_solve(int64_t n, int64_t nnz, const uint32_t *rowptr, const uint32_t *col,
                    const double *base_val, const double *base_b, int64_t n_diodes,
                    const ohm_diode_desc *diodes, int64_t n_bjts, const ohm_bjt_desc *bjts,
                    int64_t n_mosfets, const ohm_mosfet_desc *mosfets,
                    const oo_newton_params *params, double *x, int64_t *iterations,
                    double *max_diff) {
  double *xo = (double *)xmalloc(sizeof(double) * (n + 1));
  double *val = (double *)xmalloc(sizeof(double) * (nnz + 1));
  double *b = (double *)xmalloc(sizeof(double) * (n + 1));
  for (int64_t i = 0; i < n; i++) x[i] = params->initial_guess ? params->initial_guess[i] : 0.0;
  int rc = OHM_OK;
  *iterations = 0;
  *max_diff = 0.0;
  for (int64_t it = 0; it < params->max_iterations; it++) {
    memcpy(xo, x, sizeof(double) * n);
    oo_newton_assemble(n, nnz, base_val, base_b, n_diodes, diodes, n_bjts, bjts, n_mosfets,
                       mosfets, x, val, b);
    for (int64_t k = 0; k < nnz; k++)
      if (!isfinite(val[k])) {
        rc = fail(OHM_ERR_NEWTON_NUMERICAL,
                  "Newton solver encountered numerical error at iteration %lld", (long long)it);
        goto out;
      }
    rc = params->linear_mode == 1 ? oo_dense_solve(n, n, n, rowptr, col, val, b, x, 0)
                                  : oo_sparse_lu_solve(n, rowptr, col, val, b, x, 0);
    if (rc != OHM_OK) goto out;
    oo_newton_limit(n, n_diodes, diodes, n_bjts, bjts, n_mosfets, mosfets, xo, x);
    double md = 0.0;
    int bad = 0;
    for (int64_t i = 0; i < n; i++) {
      if (!isfinite(x[i])) bad = 1;
      double dd = fabs(x[i] - xo[i]);
      if (dd > md) md = dd;
    }
    *iterations = it + 1;
    *max_diff = md;
    if (bad) {
      rc = fail(OHM_ERR_NEWTON_NUMERICAL,
                "Newton solver encountered numerical error at iteration %lld", (long long)it);
      goto out;
    }
    if (md < params->abs_tol) goto out;
  }
  rc = fail(OHM_ERR_NEWTON_NOT_CONVERGED,
            "Newton solver did not converge after %lld iterations (max residual: %.2e)",
            (long long)params->max_iterations, *max_diff);
out:
  free(xo);
  free(val);
  free(b);
  return rc;
}

/* eval_pulse — src/analysis/transient_source.rs:39-71. */
double oo_eval_pulse(double v1, double v2, double td, double tr, double tf, double pw, double per,
                     double t) {
  if (t < td) return v1;
  double t_rel = (per > 0.0 && per < 1.79769313486231570815e308) ? fmod(t - td, per) : t - td;
  if (t_rel < tr) return tr > 0.0 ? v1 + (v2 - v1) * t_rel / tr : v2;
  if (t_rel < tr + pw) return v2;
  if (t_rel < tr + pw + tf) return tf > 0.0 ? v2 + (v1 - v2) * (t_rel - tr - pw) / tf : v1;
  return v1;
}

/* ------------------------------------------------------------------------ */
/* Transient — src/analysis/transient.rs:25-362, transient_source.rs:8-116.  */
/* ------------------------------------------------------------------------ */
double oo_source_value(const ohm_tran_source *s, const double *pwl, double t) {
  const double *p = s->param;
  switch (s->kind) {
    case OHM_SRC_PULSE:
      return oo_eval_pulse(p[0], p[1], p[2], p[3], p[4], p[5], p[6], t);
    case OHM_SRC_SIN: { /* eval_sin :73-83 */
      if (t < p[3]) return p[0];
      double dt = t - p[3];
      double env = p[4] != 0.0 ? exp(-dt * p[4]) : 1.0;
      return p[0] + p[1] * sin(2.0 * 3.14159265358979323846 * p[2] * dt) * env;
    }
    case OHM_SRC_EXP: { /* eval_exp :106-116 */
      if (t < p[2]) return p[0];
      double rise = (p[1] - p[0]) * (1.0 - exp(-(t - p[2]) / p[3]));
      if (t < p[4]) return p[0] + rise;
      double fall = (p[0] - p[1]) * (1.0 - exp(-(t - p[4]) / p[5]));
      return p[0] + rise + fall;
    }
    case OHM_SRC_PWL: { /* eval_pwl :85-104 */
      const double *q = pwl + 2 * s->pwl_offset;
      int64_t m = s->pwl_count;
      if (m == 0) return 0.0;
      if (t <= q[0]) return q[1];
      if (t >= q[2 * (m - 1)]) return q[2 * (m - 1) + 1];
      for (int64_t i = 1; i < m; i++)
        if (t <= q[2 * i]) {
          double frac = (t - q[2 * i - 2]) / (q[2 * i] - q[2 * i - 2]);
          return q[2 * i - 1] + (q[2 * i + 1] - q[2 * i - 1]) * frac;
        }
      return q[2 * (m - 1) + 1];
    }
    default:
      return s->dc;
  }
}

/* evaluate_b_at_time :265-280 — starts from zeros, sources in order */
static void tr_b_at(int64_t n, int64_t n_src, const ohm_tran_source *src, const double *pwl,
                    double t, double *b) {
  for (int64_t i = 0; i < n; i++) b[i] = 0.0;
  for (int64_t k = 0; k < n_src; k++) {
    double v = oo_source_value(&src[k], pwl, t);
    for (int j = 0; j < src[k].n_idx; j++) b[src[k].idx[j]] += src[k].sign[j] * v;
  }
}

/* form_companion_matrix :285-323 (A on G's pattern) */
static void tr_companion(int64_t n, const uint32_t *grp, const uint32_t *gcl, const double *gv,
                         const uint32_t *crp, const uint32_t *ccl, const double *cv, double h,
                         double alpha, double *a) {
  double factor = alpha / h;
  for (uint32_t k = 0; k < grp[n]; k++) a[k] = gv[k];
  for (int64_t r = 0; r < n; r++)
    for (uint32_t q = crp[r]; q < crp[r + 1]; q++) {
      if (cv[q] == 0.0) continue;
      for (uint32_t k = grp[r]; k < grp[r + 1]; k++)
        if (gcl[k] == ccl[q]) {
          a[k] += factor * cv[q];
          break;
        }
    }
}

int oo_transient(int64_t n, const uint32_t *grp, const uint32_t *gcl, const double *gv,
                 const uint32_t *crp, const uint32_t *ccl, const double *cv, int64_t n_src,
                 const ohm_tran_source *src, const double *pwl, int method, double tstep,
                 double tstop, double tstart, const double *x0, int64_t n_probes,
                 const uint32_t *probes, int64_t max_rec, double *times, double *waves,
                 int64_t *n_rec, double *x_final, int64_t *accepted, int64_t *rejected) {
  const double h_min = tstep / 10000.0, h_max = tstep, lte_tol = 1e-3;
  const int max_fail = 10;
  double *x = (double *)xcalloc(n > 0 ? n : 1, sizeof(double));
  double *b_next = (double *)xcalloc(n > 0 ? n : 1, sizeof(double));
  double *b_prev = (double *)xcalloc(n > 0 ? n : 1, sizeof(double));
  double *rhs = (double *)xcalloc(n > 0 ? n : 1, sizeof(double));
  double *cx = (double *)xcalloc(n > 0 ? n : 1, sizeof(double));
  double *gx = (double *)xcalloc(n > 0 ? n : 1, sizeof(double));
  double *xt = (double *)xcalloc(n > 0 ? n : 1, sizeof(double));
  double *xb = (double *)xcalloc(n > 0 ? n : 1, sizeof(double));
  double *a = (double *)xcalloc(grp[n] > 0 ? grp[n] : 1, sizeof(double));
  if (x0)
    for (int64_t i = 0; i < n; i++) x[i] = x0[i];
  int64_t rec = 0, acc = 0, rej = 0;
  int rc = OHM_OK;
#define RECORD(tt)                                                          \
  do {                                                                      \
    if (rec < max_rec) {                                                    \
      times[rec] = (tt);                                                    \
      for (int64_t j = 0; j < n_probes; j++) waves[rec * n_probes + j] = x[probes[j]]; \
      rec++;                                                                \
    }                                                                       \
  } while (0)
  if (tstart <= 0.0) RECORD(0.0);
  double t = 0.0, h = tstep < h_max ? tstep : h_max;
  tr_b_at(n, n_src, src, pwl, 0.0, b_prev);
  int use_be = 1, fails = 0;
  while (t < tstop) {
    if (t + h > tstop) h = tstop - t;
    if (h < h_min * 0.5) break;
    double t_next = t + h;
    tr_b_at(n, n_src, src, pwl, t_next, b_next);
    oo_spmv(n, crp, ccl, cv, x, cx, 0);
    if (use_be || method == OHM_TRAN_FIXED_BE) {
      tr_companion(n, grp, gcl, gv, crp, ccl, cv, h, 1.0, a);
      for (int64_t i = 0; i < n; i++) rhs[i] = b_next[i] + (1.0 / h) * cx[i];
      if ((rc = oo_sparse_lu_solve(n, grp, gcl, a, rhs, xb, 0)) != OHM_OK) break;
      for (int64_t i = 0; i < n; i++) xt[i] = xb[i];
    } else {
      oo_spmv(n, grp, gcl, gv, x, gx, 0);
      tr_companion(n, grp, gcl, gv, crp, ccl, cv, h, 2.0, a);
      double f = 2.0 / h;
      for (int64_t i = 0; i < n; i++) rhs[i] = b_next[i] + f * cx[i] - gx[i] + b_prev[i];
      if ((rc = oo_sparse_lu_solve(n, grp, gcl, a, rhs, xt, 0)) != OHM_OK) break;
      tr_companion(n, grp, gcl, gv, crp, ccl, cv, h, 1.0, a);
      for (int64_t i = 0; i < n; i++) rhs[i] = b_next[i] + (1.0 / h) * cx[i];
      if ((rc = oo_sparse_lu_solve(n, grp, gcl, a, rhs, xb, 0)) != OHM_OK) break;
      double md = 0.0;
      for (int64_t i = 0; i < n; i++) md = fmax(md, fabs(xt[i] - xb[i]));
      double lte = (2.0 / 3.0) * md;
      double hn = h * fmin(2.0, fmax(0.5, 0.9 * sqrt(lte_tol / lte)));
      if (lte > lte_tol) {
        h = fmax(hn, h_min);
        fails++;
        if (fails >= max_fail && h <= h_min * 1.01) {
          rc = fail(OHM_ERR_TIMESTEP, "Timestep error: simulation failed: %d consecutive failures", fails);
          break;
        }
        rej++;
        use_be = 1;
        continue;
      }
      h = fmin(h_max, fmax(h_min, hn));
    }
    if (use_be && method == OHM_TRAN_ADAPTIVE) use_be = 0;
    fails = 0;
    acc++;
    for (int64_t i = 0; i < n; i++) {
      x[i] = xt[i];
      b_prev[i] = b_next[i];
    }
    t = t_next;
    if (t >= tstart) RECORD(t);
  }
#undef RECORD
  if (x_final)
    for (int64_t i = 0; i < n; i++) x_final[i] = x[i];
  *n_rec = rec;
  if (accepted) *accepted = acc;
  if (rejected) *rejected = rej;
  free(x); free(b_next); free(b_prev); free(rhs); free(cx); free(gx); free(xt); free(xb); free(a);
  return rc;
}
