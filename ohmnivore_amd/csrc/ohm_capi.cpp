// ohm_capi.cpp — extern "C" entry points of include/ohm_solver.h: backend
// lifecycle, the SolverBackend buffer/matrix operations, preconditioners,
// BiCGSTAB and LinearSolver::solve_real.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "ohm_internal.h"
#include "ohm_precond.h"

namespace ohm {

static thread_local std::string g_last_error;
static std::vector<const ohm_backend *> g_live_backends;

bool backend_alive(const ohm_backend *b) {
  for (const ohm_backend *q : g_live_backends)
    if (q == b) return true;
  return false;
}

void set_error(const std::string &msg) { g_last_error = msg; }

ohm_status fail(ohm_status code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

ohm_status hip_fail(hipError_t e, const char *what, const char *file, int line) {
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
    return fail(OHM_ERR_NO_DEVICE, "Solve error: no GPU adapter (%s: %s)", what,
                hipGetErrorString(e));
  return fail(OHM_ERR_SOLVE, "Solve error: HIP call %s failed: %s (%s:%d)", what,
              hipGetErrorString(e), file, line);
}

}  // namespace ohm

using namespace ohm;

struct ohm_precond {
  ohm_backend *owner = nullptr;
  PrecondDev dev;
};

extern "C" {

const char *ohm_last_error(void) { return g_last_error.c_str(); }
const char *ohm_version(void) { return "ohmnivore-amd 0.1 (gfx950)"; }

// ------------------------------------------------------------ backend ----
ohm_status ohm_backend_create(int device, ohm_backend **out) {
  if (!out) return fail(OHM_ERR_INVALID, "null output handle");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= device || device < 0)
    return fail(OHM_ERR_NO_DEVICE, "Solve error: no GPU adapter (device %d of %d)", device,
                count);
  OHM_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  OHM_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(OHM_ERR_NO_DEVICE, "Solve error: no GPU adapter (device %d is %s, need gfx950)",
                device, prop.gcnArchName);
  ohm_backend *b = new ohm_backend();
  b->device = device;
  b->num_cu = prop.multiProcessorCount;
  OHM_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  OHM_HIP(hipMalloc(&b->d_counter, sizeof(unsigned) * 4));
  OHM_HIP(hipMemset(b->d_counter, 0, sizeof(unsigned) * 4));
  OHM_HIP(hipMalloc(&b->d_scalars, sizeof(double) * 16));
  OHM_HIP(hipHostMalloc((void **)&b->h_pinned, sizeof(double) * 16, hipHostMallocDefault));
  OHM_HIP(hipEventCreate(&b->ev0));
  OHM_HIP(hipEventCreate(&b->ev1));
  ohm_status s = ensure_partials(b, (size_t)kMaxReduceBlocks * 4);
  if (s == OHM_OK) s = static_scratch_init(b);
  if (s != OHM_OK) return s;
  *out = b;
  g_live_backends.push_back(b);
  return OHM_OK;
}

void ohm_backend_destroy(ohm_backend *b) {
  if (!b) return;
  for (size_t i = 0; i < g_live_backends.size(); i++)
    if (g_live_backends[i] == b) {
      g_live_backends.erase(g_live_backends.begin() + i);
      break;
    }
  (void)hipStreamSynchronize(b->stream);
  graph_cache_clear(b);
  if (b->d_partials) (void)hipFree(b->d_partials);
  if (b->d_counter) (void)hipFree(b->d_counter);
  if (b->d_scalars) (void)hipFree(b->d_scalars);
  if (b->h_pinned) (void)hipHostFree(b->h_pinned);
  if (b->ac_ws) (void)hipFree(b->ac_ws);
  if (b->dense_ws) (void)hipFree(b->dense_ws);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

ohm_status ohm_backend_sync(ohm_backend *b) {
  if (!b) return fail(OHM_ERR_INVALID, "null backend");
  OHM_HIP(hipStreamSynchronize(b->stream));
  return OHM_OK;
}

ohm_status ohm_backend_set_graphs(ohm_backend *b, int enable) {
  if (!b) return fail(OHM_ERR_INVALID, "null backend");
  b->use_graphs = enable != 0;
  if (!b->use_graphs) graph_cache_clear(b);
  return OHM_OK;
}

ohm_status ohm_backend_counters(const ohm_backend *b, int64_t *dispatches, int64_t *readbacks) {
  if (!b) return fail(OHM_ERR_INVALID, "null backend");
  if (dispatches) *dispatches = b->dispatches;
  if (readbacks) *readbacks = b->readbacks;
  return OHM_OK;
}

// ------------------------------------------------------------ buffers ----
ohm_status ohm_buffer_new(ohm_backend *b, size_t n, ohm_buffer **out) {
  if (!b || !out) return fail(OHM_ERR_INVALID, "null handle");
  ohm_buffer *buf = new ohm_buffer();
  buf->owner = b;
  buf->n = n;
  hipError_t e = hipMalloc(&buf->d, sizeof(double) * (n ? n : 1));
  if (e != hipSuccess) {
    delete buf;
    return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
  }
  OHM_HIP(hipMemsetAsync(buf->d, 0, sizeof(double) * (n ? n : 1), b->stream));
  OHM_HIP(hipStreamSynchronize(b->stream));
  *out = buf;
  return OHM_OK;
}

void ohm_buffer_free(ohm_buffer *buf) {
  if (!buf) return;
  if (buf->d) (void)hipFree(buf->d);
  delete buf;
}

size_t ohm_buffer_len(const ohm_buffer *buf) { return buf ? buf->n : 0; }

ohm_status ohm_upload_vec(ohm_backend *b, const double *data, size_t n, ohm_buffer *buf) {
  if (!b || !buf) return fail(OHM_ERR_INVALID, "null handle");
  if (n > buf->n)
    return fail(OHM_ERR_SOLVE, "Solve error: dimension mismatch: upload of %zu into buffer of %zu",
                n, buf->n);
  if (n) OHM_HIP(hipMemcpyAsync(buf->d, data, sizeof(double) * n, hipMemcpyHostToDevice, b->stream));
  OHM_HIP(hipStreamSynchronize(b->stream));
  return OHM_OK;
}

ohm_status ohm_download_vec(ohm_backend *b, const ohm_buffer *buf, double *out, size_t n) {
  if (!b || !buf) return fail(OHM_ERR_INVALID, "null handle");
  if (n > buf->n)
    return fail(OHM_ERR_SOLVE, "Solve error: dimension mismatch: download of %zu from buffer of %zu",
                n, buf->n);
  if (n) OHM_HIP(hipMemcpyAsync(out, buf->d, sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
  OHM_HIP(hipStreamSynchronize(b->stream));
  b->readbacks++;
  return OHM_OK;
}

// ----------------------------------------------------------- matrices ----
static ohm_status validate_csr(const uint32_t *col, const uint32_t *rowptr, size_t n, size_t nnz) {
  if (n > 0x7fffffffu || nnz > 0xffffffffu)
    return fail(OHM_ERR_SOLVE, "Solve error: matrix too large for 32-bit CSR (%zu rows, %zu nnz)",
                n, nnz);
  if (rowptr[0] != 0 || rowptr[n] != nnz)
    return fail(OHM_ERR_SOLVE, "Solve error: dimension mismatch: row_pointers[n]=%u, nnz=%zu",
                rowptr[n], nnz);
  for (size_t r = 0; r < n; r++) {
    if (rowptr[r + 1] < rowptr[r])
      return fail(OHM_ERR_SOLVE, "Solve error: row_pointers not monotone at row %zu", r);
    for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++)
      if (col[k] >= n)
        return fail(OHM_ERR_SOLVE, "Solve error: column index %u out of range in row %zu", col[k],
                    r);
  }
  return OHM_OK;
}

ohm_status ohm_upload_matrix(ohm_backend *b, const double *values, const uint32_t *col_indices,
                             const uint32_t *row_pointers, size_t n, size_t nnz,
                             ohm_matrix **out) {
  if (!b || !out || !row_pointers || (nnz && (!col_indices || !values)))
    return fail(OHM_ERR_INVALID, "null handle");
  OHM_TRY(validate_csr(col_indices, row_pointers, n, nnz));
  ohm_matrix *m = new ohm_matrix();
  m->owner = b;
  ohm_status s = csr_upload(b, m->csr, values, col_indices, row_pointers, (int)n, (int64_t)nnz);
  if (s != OHM_OK) {
    csr_free(m->csr);
    delete m;
    return s;
  }
  *out = m;
  return OHM_OK;
}

void ohm_matrix_free(ohm_matrix *m) {
  if (!m) return;
  if (backend_alive(m->owner)) graph_cache_clear(m->owner);
  csr_free(m->csr);
  delete m;
}

ohm_status ohm_matrix_set_values(ohm_backend *b, ohm_matrix *m, const double *values) {
  if (!b || !m) return fail(OHM_ERR_INVALID, "null handle");
  if (m->csr.nnz)
    OHM_HIP(hipMemcpyAsync(m->csr.val, values, sizeof(double) * m->csr.nnz, hipMemcpyHostToDevice,
                           b->stream));
  if (m->csr.val32) OHM_TRY(csr_ensure_f32(b, m->csr));
  OHM_HIP(hipStreamSynchronize(b->stream));
  return OHM_OK;
}

static ohm_status check_len(const ohm_buffer *x, size_t n, const char *what) {
  if (!x) return fail(OHM_ERR_INVALID, "null buffer");
  if (x->n < n)
    return fail(OHM_ERR_SOLVE, "Solve error: dimension mismatch: %s has %zu entries, need %zu", what,
                x->n, n);
  return OHM_OK;
}

ohm_status ohm_spmv(ohm_backend *b, const ohm_matrix *a, const ohm_buffer *x, ohm_buffer *y) {
  if (!b || !a) return fail(OHM_ERR_INVALID, "null handle");
  OHM_TRY(check_len(x, a->csr.n, "x"));
  OHM_TRY(check_len(y, a->csr.n, "y"));
  launch_spmv(b, a->csr, x->d, y->d, nullptr, 0);
  OHM_HIP(hipGetLastError());
  OHM_HIP(hipStreamSynchronize(b->stream));
  return OHM_OK;
}

ohm_status ohm_backend_set_autotune(ohm_backend *b, int enable) {
  if (!b) return fail(OHM_ERR_INVALID, "null backend");
  b->autotune = enable != 0;
  return OHM_OK;
}

ohm_status ohm_matrix_autotune(ohm_backend *b, ohm_matrix *m, int32_t *rows_per_block,
                               int32_t *nnz_per_block) {
  if (!b || !m) return fail(OHM_ERR_INVALID, "null handle");
  int r = 0, z = 0;
  OHM_TRY(csr_autotune(b, m->csr, &r, &z));
  if (rows_per_block) *rows_per_block = r;
  if (nnz_per_block) *nnz_per_block = z;
  return OHM_OK;
}

ohm_status ohm_stream_timed(ohm_backend *b, int64_t bytes, int mode, int reps, double *avg_ms) {
  if (!b || !avg_ms || reps <= 0 || bytes <= 0 || (mode != 0 && mode != 1))
    return fail(OHM_ERR_INVALID, "bad argument");
  double ms = 0.0;
  OHM_TRY(stream_timed(b, bytes, mode, reps, ms));
  *avg_ms = ms;
  return OHM_OK;
}

ohm_status ohm_spmv_timed(ohm_backend *b, const ohm_matrix *a, const ohm_buffer *x, ohm_buffer *y,
                          int reps, int precision, double *avg_ms) {
  if (!b || !a || !avg_ms || reps <= 0) return fail(OHM_ERR_INVALID, "bad argument");
  OHM_TRY(check_len(x, a->csr.n, "x"));
  OHM_TRY(check_len(y, a->csr.n, "y"));
  DevCsr &m = const_cast<ohm_matrix *>(a)->csr;
  float *x32 = nullptr, *y32 = nullptr;
  if (precision == OHM_PRECISION_FP32_IR) {
    OHM_TRY(csr_ensure_f32(b, m));
    OHM_HIP(hipMalloc(&x32, sizeof(float) * (m.n ? m.n : 1)));
    OHM_HIP(hipMalloc(&y32, sizeof(float) * (m.n ? m.n : 1)));
    OHM_HIP(hipMemsetAsync(x32, 0, sizeof(float) * (m.n ? m.n : 1), b->stream));
  }
  auto one = [&]() {
    if (x32)
      launch_spmv_f32(b, m, x32, y32, nullptr, 0);
    else
      launch_spmv(b, m, x->d, y->d, nullptr, 0);
  };
  one();
  OHM_HIP(hipEventRecord(b->ev0, b->stream));
  for (int k = 0; k < reps; k++) one();
  OHM_HIP(hipEventRecord(b->ev1, b->stream));
  OHM_HIP(hipEventSynchronize(b->ev1));
  OHM_HIP(hipGetLastError());
  float ms = 0.f;
  OHM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  *avg_ms = (double)ms / reps;
  if (x32) (void)hipFree(x32);
  if (y32) (void)hipFree(y32);
  return OHM_OK;
}

ohm_status ohm_spmv_transpose(ohm_backend *b, const ohm_matrix *a, const ohm_buffer *x,
                              ohm_buffer *y) {
  if (!b || !a) return fail(OHM_ERR_INVALID, "null handle");
  OHM_TRY(check_len(x, a->csr.n, "x"));
  OHM_TRY(check_len(y, a->csr.n, "y"));
  OHM_TRY(csr_ensure_transpose(b, const_cast<ohm_matrix *>(a)->csr));
  launch_spmv_t(b, a->csr, x->d, y->d);
  OHM_HIP(hipGetLastError());
  OHM_HIP(hipStreamSynchronize(b->stream));
  return OHM_OK;
}

ohm_status ohm_dot_n(ohm_backend *b, const ohm_buffer *x, const ohm_buffer *y, size_t n,
                     double *out) {
  if (!b || !out) return fail(OHM_ERR_INVALID, "null handle");
  OHM_TRY(check_len(x, n, "x"));
  OHM_TRY(check_len(y, n, "y"));
  if (n == 0) {
    *out = 0.0;
    return OHM_OK;
  }
  launch_dot(b, (int)n, x->d, y->d, 0);
  OHM_HIP(hipGetLastError());
  OHM_HIP(hipMemcpyAsync(b->h_pinned, b->d_scalars, sizeof(double), hipMemcpyDeviceToHost,
                         b->stream));
  OHM_HIP(hipStreamSynchronize(b->stream));
  b->readbacks++;
  *out = b->h_pinned[0];
  return OHM_OK;
}

ohm_status ohm_dot(ohm_backend *b, const ohm_buffer *x, const ohm_buffer *y, double *out) {
  if (!x || !y) return fail(OHM_ERR_INVALID, "null buffer");
  return ohm_dot_n(b, x, y, x->n < y->n ? x->n : y->n, out);
}

ohm_status ohm_axpy(ohm_backend *b, double alpha, const ohm_buffer *x, ohm_buffer *y) {
  if (!b || !x || !y) return fail(OHM_ERR_INVALID, "null handle");
  OHM_TRY(check_len(x, y->n, "x"));
  launch_axpy(b, (int)y->n, alpha, x->d, y->d);
  OHM_HIP(hipGetLastError());
  OHM_HIP(hipStreamSynchronize(b->stream));
  return OHM_OK;
}

ohm_status ohm_copy(ohm_backend *b, const ohm_buffer *src, ohm_buffer *dst) {
  if (!b || !src || !dst) return fail(OHM_ERR_INVALID, "null handle");
  OHM_TRY(check_len(src, dst->n, "src"));
  launch_copy(b, (int)dst->n, src->d, dst->d);
  OHM_HIP(hipGetLastError());
  OHM_HIP(hipStreamSynchronize(b->stream));
  return OHM_OK;
}

ohm_status ohm_scale(ohm_backend *b, double alpha, ohm_buffer *x) {
  if (!b || !x) return fail(OHM_ERR_INVALID, "null handle");
  launch_scale(b, (int)x->n, alpha, x->d);
  OHM_HIP(hipGetLastError());
  OHM_HIP(hipStreamSynchronize(b->stream));
  return OHM_OK;
}

// ------------------------------------------------------ preconditioners --
ohm_status ohm_precond_create(ohm_backend *b, const ohm_matrix *a, int kind, int level,
                              ohm_precond **out) {
  if (!b || !a || !out) return fail(OHM_ERR_INVALID, "null handle");
  ohm_precond *p = new ohm_precond();
  p->owner = b;
  ohm_status s = precond_build(b, a->csr, kind, level, p->dev);
  if (s == OHM_OK) {
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) s = hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
  }
  if (s != OHM_OK) {
    precond_free(p->dev);
    delete p;
    return s;
  }
  *out = p;
  return OHM_OK;
}

ohm_status ohm_precond_update(ohm_backend *b, ohm_precond *p, const ohm_matrix *a) {
  if (!b || !p || !a) return fail(OHM_ERR_INVALID, "null handle");
  OHM_TRY(precond_refresh(b, a->csr, p->dev));
  OHM_HIP(hipStreamSynchronize(b->stream));
  return OHM_OK;
}

ohm_status ohm_precond_apply(ohm_backend *b, const ohm_precond *p, const ohm_buffer *in,
                             ohm_buffer *out) {
  if (!b || !p) return fail(OHM_ERR_INVALID, "null handle");
  OHM_TRY(check_len(in, p->dev.n, "in"));
  OHM_TRY(check_len(out, p->dev.n, "out"));
  double *tmp = nullptr;
  OHM_HIP(hipMalloc(&tmp, sizeof(double) * (p->dev.n ? p->dev.n : 1)));
  precond_apply(b, p->dev, in->d, out->d, tmp, nullptr, 0);
  hipError_t e = hipStreamSynchronize(b->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return hip_fail(e, "precond_apply", __FILE__, __LINE__);
  return OHM_OK;
}

int ohm_precond_kind(const ohm_precond *p) {
  if (!p) return -1;
  return p->dev.kind == PK_JACOBI ? OHM_PREC_JACOBI
                                  : (p->dev.kind == PK_ISAI ? OHM_PREC_ISAI : OHM_PREC_NONE);
}

ohm_status ohm_precond_set_ilu_sweeps(ohm_precond *p, int sweeps) {
  if (!p || sweeps < 0) return fail(OHM_ERR_INVALID, "bad preconditioner or sweep count");
  p->dev.ilu_sweeps = sweeps;
  isai_reset(p->dev);
  return OHM_OK;
}

void ohm_precond_free(ohm_precond *p) {
  if (!p) return;
  if (backend_alive(p->owner)) graph_cache_clear(p->owner);
  precond_free(p->dev);
  delete p;
}

// ----------------------------------------------------------- BiCGSTAB ----
void ohm_bicgstab_default_params(ohm_bicgstab_params *p) {
  if (!p) return;
  p->max_iterations = 10000;
  p->tolerance = 1e-5;
  p->shadow = OHM_SHADOW_HASHED;
  p->warm_start = 0;
  p->precision = OHM_PRECISION_FP64;
  p->check_interval = 0;
  p->inner_tolerance = 1e-4;
  p->max_refinements = 30;
  p->breakdown = OHM_BREAKDOWN_SCALED;
  p->tol_mode = OHM_TOL_RHS;
  p->tol_floor = 1e-15;
}

static ohm_status run_bicgstab(ohm_backend *b, DevCsr &a, PrecondDev &pd, const double *rhs,
                               double *x, const ohm_bicgstab_params *params,
                               ohm_bicgstab_result *result, bool transpose) {
  ohm_bicgstab_params prm;
  if (params)
    prm = *params;
  else
    ohm_bicgstab_default_params(&prm);
  BicgOpts o;
  o.max_iters = prm.max_iterations;
  o.tol = prm.tolerance;
  o.shadow = prm.shadow;
  o.warm = prm.warm_start != 0;
  o.check_interval = prm.check_interval;
  o.transpose = transpose;
  o.scaled_breakdown = prm.breakdown != OHM_BREAKDOWN_ABSOLUTE;
  o.tol_mode = prm.tol_mode;
  o.tol_floor = prm.tol_floor;
  BicgOut out;
  int64_t refinements = 0;
  int64_t d0 = b->dispatches;
  OHM_HIP(hipEventRecord(b->ev0, b->stream));
  ohm_status s;
  if (transpose)  // adjoint systems run in FP64
    s = bicg_solve_transpose(b, a, pd, rhs, x, o, out);
  else if (prm.precision == OHM_PRECISION_FP32_IR)
    s = bicg_solve_ir(b, a, pd, rhs, x, o, prm.inner_tolerance, prm.max_refinements, out,
                      refinements);
  else
    s = bicg_solve_f64(b, a, pd, rhs, x, o, out);
  OHM_HIP(hipEventRecord(b->ev1, b->stream));
  OHM_HIP(hipEventSynchronize(b->ev1));
  if (result) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    result->iterations = out.iters;
    result->refinements = refinements;
    result->fp64_fallback = out.fp64_fallback ? 1 : 0;
    result->device_ms = ms;
    result->dispatches = b->dispatches - d0;
    result->readbacks = out.polls;
    double rel = NAN;
    if (s == OHM_OK || s == OHM_ERR_SOLVE) {
      std::string keep = g_last_error;
      DevCsr at;
      const DevCsr *ra = &a;
      if (transpose && csr_transpose_view(b, a, at) == OHM_OK) ra = &at;
      if (true_residual(b, *ra, rhs, x, rel) != OHM_OK) rel = NAN;
      g_last_error = keep;
    }
    result->residual_norm = rel;
  }
  return s;
}

ohm_status ohm_bicgstab(ohm_backend *b, const ohm_matrix *a, const ohm_buffer *rhs, ohm_buffer *x,
                        const ohm_precond *p, const ohm_bicgstab_params *params,
                        ohm_bicgstab_result *result) {
  if (!b || !a) return fail(OHM_ERR_INVALID, "null handle");
  OHM_TRY(check_len(rhs, a->csr.n, "rhs"));
  OHM_TRY(check_len(x, a->csr.n, "x"));
  PrecondDev none;
  none.n = a->csr.n;
  PrecondDev &pd = p ? const_cast<ohm_precond *>(p)->dev : none;
  return run_bicgstab(b, const_cast<ohm_matrix *>(a)->csr, pd, rhs->d, x->d, params, result,
                      false);
}

ohm_status ohm_bicgstab_noalloc(ohm_backend *b, const ohm_matrix *a, const ohm_buffer *rhs,
                                ohm_buffer *x, const ohm_precond *p,
                                const ohm_bicgstab_params *params, ohm_bicgstab_result *result) {
  if (!b || !a) return fail(OHM_ERR_INVALID, "null handle");
  OHM_TRY(check_len(rhs, a->csr.n, "rhs"));
  OHM_TRY(check_len(x, a->csr.n, "x"));
  ohm_bicgstab_params prm;
  if (params)
    prm = *params;
  else
    ohm_bicgstab_default_params(&prm);
  if (prm.precision != OHM_PRECISION_FP64)
    return fail(OHM_ERR_INVALID, "ohm_bicgstab_noalloc solves in FP64 only");
  PrecondDev none;
  none.n = a->csr.n;
  const PrecondDev &pd = p ? p->dev : none;
  BicgOpts o;
  o.max_iters = prm.max_iterations;
  o.tol = prm.tolerance;
  o.shadow = prm.shadow;
  o.warm = prm.warm_start != 0;
  o.check_interval = prm.check_interval;
  o.scaled_breakdown = prm.breakdown != OHM_BREAKDOWN_ABSOLUTE;
  o.tol_mode = prm.tol_mode;
  o.tol_floor = prm.tol_floor;
  BicgOut out;
  double rel = NAN;
  const int64_t d0 = b->dispatches;
  OHM_HIP(hipEventRecord(b->ev0, b->stream));
  ohm_status s = bicg_solve_f64_noalloc(b, a->csr, pd, rhs->d, x->d, o, out, rel);
  OHM_HIP(hipEventRecord(b->ev1, b->stream));
  OHM_HIP(hipEventSynchronize(b->ev1));
  if (result) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    result->iterations = out.iters;
    result->refinements = 0;
    result->fp64_fallback = 0;
    result->device_ms = ms;
    result->dispatches = b->dispatches - d0;
    result->readbacks = out.polls;
    result->residual_norm = (s == OHM_OK || s == OHM_ERR_SOLVE) ? rel : NAN;
  }
  return s;
}

ohm_status ohm_bicgstab_transpose(ohm_backend *b, const ohm_matrix *a, const ohm_buffer *rhs,
                                  ohm_buffer *x, const ohm_precond *p,
                                  const ohm_bicgstab_params *params, ohm_bicgstab_result *result) {
  if (!b || !a || !rhs || !x) return fail(OHM_ERR_INVALID, "null handle");
  if (a->csr.n != (int)rhs->n || a->csr.n != (int)x->n)
    return fail(OHM_ERR_SOLVE, "Solve error: dimension mismatch");
  PrecondDev none;
  none.n = a->csr.n;
  PrecondDev &pd = p ? const_cast<ohm_precond *>(p)->dev : none;
  return run_bicgstab(b, const_cast<ohm_matrix *>(a)->csr, pd, rhs->d, x->d, params, result, true);
}

ohm_status ohm_solve_real(ohm_backend *b, size_t n, size_t nnz, const double *values,
                          const uint32_t *col_indices, const uint32_t *row_pointers,
                          const double *rhs, double *x, int precond,
                          const ohm_bicgstab_params *params, ohm_bicgstab_result *result) {
  if (!b) return fail(OHM_ERR_INVALID, "null backend");
  if (n == 0) return OHM_OK;  // src/solver/gpu.rs:117-119
  ohm_matrix *a = nullptr;
  OHM_TRY(ohm_upload_matrix(b, values, col_indices, row_pointers, n, nnz, &a));
  ohm_buffer *bb = nullptr, *xx = nullptr;
  ohm_precond *p = nullptr;
  ohm_status s = ohm_buffer_new(b, n, &bb);
  if (s == OHM_OK) s = ohm_buffer_new(b, n, &xx);
  if (s == OHM_OK) s = ohm_upload_vec(b, rhs, n, bb);
  if (s == OHM_OK && params && params->warm_start) s = ohm_upload_vec(b, x, n, xx);
  if (s == OHM_OK) s = ohm_precond_create(b, a, precond, 1, &p);
  if (s == OHM_OK) s = ohm_bicgstab(b, a, bb, xx, p, params, result);
  if (s == OHM_OK) s = ohm_download_vec(b, xx, x, n);
  std::string keep = g_last_error;
  ohm_precond_free(p);
  ohm_buffer_free(bb);
  ohm_buffer_free(xx);
  ohm_matrix_free(a);
  g_last_error = keep;
  return s;
}

}  // extern "C"
