// pf_capi.hip -- extern "C" entry points of libpfpa.so (declared in include/pfpa.h).
#include <hip/hip_runtime.h>

#include <string>

#include "pfpa.h"
#include "pf_internal.h"
#include "pf_thermo.h"

struct pf_ctx {
  pf::Scratch scratch;
  pf::TileConfig tile;
  int device;
  double xoff[2];  // composition offset of FP32-stored arrays (pf_ctx_set_storage_offset)
};

namespace {

int fail(int code, const std::string& msg) {
  pf::set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  return fail(PF_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

int check_grid(int nx, int ny) {
  if (nx < 3 || ny < 3) return fail(PF_ERR_ARG, "grid must have at least one interior cell (nx, ny >= 3)");
  if ((long long)nx * (long long)ny >= (1LL << 31))
    return fail(PF_ERR_ARG, "grid too large for 32-bit cell indices");
  return PF_OK;
}

// The composition offset the arrays of storage type T hold on this context:
// FP32 storage may carry one, FP64 storage never does.
template <typename T>
const double* storage_offset(const pf_ctx* ctx) {
  return sizeof(T) == sizeof(float) ? ctx->xoff : nullptr;
}

int check_ptrs(const void* a, const void* b, const void* c, const void* d) {
  if (!a || !b || !c || !d) return fail(PF_ERR_ARG, "null field pointer");
  return PF_OK;
}

// A step reads the four old fields while it writes the four new ones, so no
// new field may overlap an old field or another new field (each array holds
// nx * ny cells).
template <typename T>
int check_disjoint(int nx, int ny, const T* i0, const T* i1, const T* i2, const T* i3,
                   const T* o0, const T* o1, const T* o2, const T* o3) {
  const uintptr_t n = (uintptr_t)nx * (uintptr_t)ny * sizeof(T);
  const uintptr_t in[4] = {(uintptr_t)i0, (uintptr_t)i1, (uintptr_t)i2, (uintptr_t)i3};
  const uintptr_t out[4] = {(uintptr_t)o0, (uintptr_t)o1, (uintptr_t)o2, (uintptr_t)o3};
  auto overlap = [n](uintptr_t a, uintptr_t b) { return a < b + n && b < a + n; };
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 4; j++)
      if (overlap(out[i], in[j]))
        return fail(PF_ERR_ARG,
                    "step cannot run in place: the new field buffers must not overlap the old ones");
    for (int j = 0; j < i; j++)
      if (overlap(out[i], out[j]))
        return fail(PF_ERR_ARG, "the four new field buffers must not overlap each other");
  }
  return PF_OK;
}

template <typename T>
int step_t(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm, const T* iCr,
           const T* iNb, const T* ipd, const T* ipl, T* oCr, T* oNb, T* opd, T* opl,
           const pf_nuc_params* nuc, void* stream) {
  if (!ctx || !prm) return fail(PF_ERR_ARG, "null context or parameters");
  if (int r = check_grid(nx, ny)) return r;
  if (int r = check_ptrs(iCr, iNb, ipd, ipl)) return r;
  if (int r = check_ptrs(oCr, oNb, opd, opl)) return r;
  if (bc != PF_BC_NOFLUX && bc != PF_BC_PERIODIC) return fail(PF_ERR_ARG, "unknown boundary mode");
  if (int r = check_disjoint<T>(nx, ny, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl)) return r;
  const double* xoff = storage_offset<T>(ctx);
  const pf::StepParams P = pf::make_step(prm->dt, prm->DCr, prm->DNb, prm->kappa, prm->omega,
                                         prm->Lmob, prm->alpha, prm->mask, xoff);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  if (nuc) {
    pf::NucParams N =
        pf::make_nuc(nuc->sigma_del, nuc->sigma_lav, nuc->unit_a, nuc->dx, nuc->dy, nuc->dz,
                     nuc->dt, nuc->seed, nuc->call, nuc->instance, nx, ny);
    N.xoff[0] = P.xoff[0];
    N.xoff[1] = P.xoff[1];
    if (N.R_scan < 1.0) {
      // fused: decisions on each cell's new values, centre embeds in place,
      // wider embeds deferred to the event pass
      e = pf::launch_step<T>(nx, ny, bc, P, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, &N,
                             ctx->scratch, ctx->tile, st);
      if (e != hipSuccess) return hip_fail(e, "step kernel");
      e = pf::launch_apply_events<T>(nx, ny, oCr, oNb, opd, opl, ctx->scratch, xoff, st);
      if (e != hipSuccess) return hip_fail(e, "nucleation embed kernel");
      return PF_OK;
    }
    e = pf::launch_step<T>(nx, ny, bc, P, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, nullptr,
                           ctx->scratch, ctx->tile, st);
    if (e != hipSuccess) return hip_fail(e, "step kernel");
    e = pf::launch_nucleation<T>(nx, ny, oCr, oNb, opd, opl, N, ctx->scratch, st);
    if (e != hipSuccess) return hip_fail(e, "nucleation kernel");
    return PF_OK;
  }
  e = pf::launch_step<T>(nx, ny, bc, P, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl, nullptr,
                         ctx->scratch, ctx->tile, st);
  if (e != hipSuccess) return hip_fail(e, "step kernel");
  return PF_OK;
}

template <typename T>
int stepk_t(int k, pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm, const T* iCr,
            const T* iNb, const T* ipd, const T* ipl, T* oCr, T* oNb, T* opd, T* opl,
            void* stream) {
  if (!ctx || !prm) return fail(PF_ERR_ARG, "null context or parameters");
  if (int r = check_grid(nx, ny)) return r;
  if (nx < 5 || ny < 5)
    return fail(PF_ERR_ARG, "multi-step kernels need at least 3x3 interior cells");
  if (k == 4 && (nx < 6 || ny < 6))
    return fail(PF_ERR_ARG, "the four-step kernel needs at least 4x4 interior cells");
  if (int r = check_ptrs(iCr, iNb, ipd, ipl)) return r;
  if (int r = check_ptrs(oCr, oNb, opd, opl)) return r;
  if (bc != PF_BC_NOFLUX && bc != PF_BC_PERIODIC) return fail(PF_ERR_ARG, "unknown boundary mode");
  if (int r = check_disjoint<T>(nx, ny, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl)) return r;
  if (!pf::symmetric_five_point(prm->mask))
    return fail(PF_ERR_ARG,
                "multi-step kernels take symmetric five-point masks only (zero corners, "
                "m[0][1] = m[2][1] != 0, m[1][0] = m[1][2])");
  const pf::StepParams P = pf::make_step(prm->dt, prm->DCr, prm->DNb, prm->kappa, prm->omega,
                                         prm->Lmob, prm->alpha, prm->mask, storage_offset<T>(ctx));
  hipError_t e = pf::launch_stepk<T>(k, nx, ny, bc, P, iCr, iNb, ipd, ipl, oCr, oNb, opd, opl,
                                     ctx->tile, (hipStream_t)stream);
  if (e != hipSuccess)
    return hip_fail(e, k == 4 ? "four-step kernel" : k == 3 ? "three-step kernel" : "two-step kernel");
  return PF_OK;
}

int set_tile_k(pf_ctx* ctx, int k, int rows, int depth) {
  if (!ctx) return fail(PF_ERR_ARG, "null context");
  if (rows < 0 || rows > 1 << 20) return fail(PF_ERR_ARG, "rows out of range");
  if (depth != 0 && depth != 2 && depth != 3 && depth != 6)
    return fail(PF_ERR_ARG, "prefetch depth must be 2, 3 or 6 (0 = default)");
  (k == 4 ? ctx->tile.rows4 : k == 3 ? ctx->tile.rows3 : ctx->tile.rows2) = rows;
  (k == 4 ? ctx->tile.depth4 : k == 3 ? ctx->tile.depth3 : ctx->tile.depth2) = depth ? depth : 3;
  return PF_OK;
}

template <typename T>
int nucleation_t(pf_ctx* ctx, int nx, int ny, T* xCr, T* xNb, T* pd, T* pl,
                 const pf_nuc_params* nuc, void* stream) {
  if (!ctx || !nuc) return fail(PF_ERR_ARG, "null context or parameters");
  if (int r = check_grid(nx, ny)) return r;
  if (int r = check_ptrs(xCr, xNb, pd, pl)) return r;
  pf::NucParams N =
      pf::make_nuc(nuc->sigma_del, nuc->sigma_lav, nuc->unit_a, nuc->dx, nuc->dy, nuc->dz,
                   nuc->dt, nuc->seed, nuc->call, nuc->instance, nx, ny);
  if (const double* xoff = storage_offset<T>(ctx)) {
    N.xoff[0] = xoff[0];
    N.xoff[1] = xoff[1];
  }
  hipError_t e =
      pf::launch_nucleation<T>(nx, ny, xCr, xNb, pd, pl, N, ctx->scratch, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "nucleation kernel");
  return PF_OK;
}

template <typename T>
int summarize_t(pf_ctx* ctx, int nx, int ny, const T* xCr, const T* xNb, const T* pd,
                const T* pl, double* out5, void* stream) {
  if (!ctx || !out5) return fail(PF_ERR_ARG, "null context or output");
  if (int r = check_grid(nx, ny)) return r;
  if (int r = check_ptrs(xCr, xNb, pd, pl)) return r;
  hipError_t e = pf::launch_summarize<T>(nx, ny, xCr, xNb, pd, pl, storage_offset<T>(ctx), out5,
                                         ctx->scratch, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "summarize kernels");
  return PF_OK;
}

template <typename T>
int energy_t(pf_ctx* ctx, int nx, int ny, int bc, double dx, double dy, const double* kappa,
             const double* omega, double Omega, const T* xCr, const T* xNb, const T* pd,
             const T* pl, double* chem, double* grad, double* out3, void* stream) {
  if (!ctx || !kappa || !omega || !out3) return fail(PF_ERR_ARG, "null argument");
  if (int r = check_grid(nx, ny)) return r;
  if (int r = check_ptrs(xCr, xNb, pd, pl)) return r;
  if (bc != PF_BC_NOFLUX && bc != PF_BC_PERIODIC) return fail(PF_ERR_ARG, "unknown boundary mode");
  if (!(dx > 0.0) || !(dy > 0.0)) return fail(PF_ERR_ARG, "mesh spacing must be positive");
  pf::EnergyParams E;
  E.dx = dx;
  E.dy = dy;
  E.dV = dx * dy;
  for (int k = 0; k < 2; k++) {
    E.kappa[k] = kappa[k];
    E.omega[k] = omega[k];
  }
  E.Omega = Omega;
  const double* xoff = storage_offset<T>(ctx);
  E.xoff[0] = xoff ? xoff[0] : 0.0;
  E.xoff[1] = xoff ? xoff[1] : 0.0;
  hipError_t e = pf::launch_energy<T>(nx, ny, bc, E, xCr, xNb, pd, pl, chem, grad, out3,
                                      ctx->scratch, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "energy kernels");
  return PF_OK;
}

}  // namespace

extern "C" {

int pf_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int pf_ctx_create(pf_ctx** out) {
  if (!out) return fail(PF_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (pf_device_count() < 1) return fail(PF_ERR_NO_DEVICE, "no HIP device visible");
  pf_ctx* c = new pf_ctx();
  c->scratch = pf::Scratch{};
  c->xoff[0] = c->xoff[1] = 0.0;
  if (hipGetDevice(&c->device) != hipSuccess) c->device = 0;
  hipError_t e = pf::scratch_alloc(c->scratch);
  if (e != hipSuccess) {
    pf::scratch_free(c->scratch);
    delete c;
    return hip_fail(e, "scratch allocation");
  }
  *out = c;
  return PF_OK;
}

void pf_ctx_destroy(pf_ctx* ctx) {
  if (!ctx) return;
  pf::scratch_free(ctx->scratch);
  delete ctx;
}

const char* pf_last_error(void) { return pf::last_error(); }

int pf_ctx_set_tile(pf_ctx* ctx, int rows, int xcd_remap, int depth) {
  if (!ctx) return fail(PF_ERR_ARG, "null context");
  if (rows < 0 || rows > 1 << 20) return fail(PF_ERR_ARG, "rows out of range");
  if (depth != 0 && depth != 2 && depth != 3 && depth != 6)
    return fail(PF_ERR_ARG, "prefetch depth must be 2, 3 or 6 (0 = default)");
  ctx->tile.rows = rows;
  ctx->tile.remap = xcd_remap ? 1 : 0;
  ctx->tile.depth = depth ? depth : 3;
  return PF_OK;
}

int pf_default_rows(int nx, int ny) { return pf::default_rows(nx, ny); }

int pf_ctx_set_storage_offset(pf_ctx* ctx, double x_cr, double x_nb) {
  if (!ctx) return fail(PF_ERR_ARG, "null context");
  if (!(x_cr - x_cr == 0.0) || !(x_nb - x_nb == 0.0))
    return fail(PF_ERR_ARG, "storage offsets must be finite");
  ctx->xoff[0] = x_cr;
  ctx->xoff[1] = x_nb;
  return PF_OK;
}

int pf_step_f64(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm, const double* a,
                const double* b, const double* c, const double* d, double* oa, double* ob,
                double* oc, double* od, const pf_nuc_params* nuc, void* stream) {
  return step_t<double>(ctx, nx, ny, bc, prm, a, b, c, d, oa, ob, oc, od, nuc, stream);
}

int pf_step_f32(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm, const float* a,
                const float* b, const float* c, const float* d, float* oa, float* ob, float* oc,
                float* od, const pf_nuc_params* nuc, void* stream) {
  return step_t<float>(ctx, nx, ny, bc, prm, a, b, c, d, oa, ob, oc, od, nuc, stream);
}

int pf_ctx_set_tile2(pf_ctx* ctx, int rows, int depth) { return set_tile_k(ctx, 2, rows, depth); }
int pf_ctx_set_tile3(pf_ctx* ctx, int rows, int depth) { return set_tile_k(ctx, 3, rows, depth); }
int pf_ctx_set_tile4(pf_ctx* ctx, int rows, int depth) { return set_tile_k(ctx, 4, rows, depth); }

int pf_default_rows2(int nx, int ny, int storage) {
  return pf::default_rows_k(2, nx, ny, storage == PF_STORAGE_F32 ? 4 : 8);
}
int pf_default_rows3(int nx, int ny, int storage) {
  return pf::default_rows_k(3, nx, ny, storage == PF_STORAGE_F32 ? 4 : 8);
}
int pf_default_rows4(int nx, int ny, int storage) {
  return pf::default_rows_k(4, nx, ny, storage == PF_STORAGE_F32 ? 4 : 8);
}

int pf_step2_f64(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm, const double* a,
                 const double* b, const double* c, const double* d, double* oa, double* ob,
                 double* oc, double* od, void* stream) {
  return stepk_t<double>(2, ctx, nx, ny, bc, prm, a, b, c, d, oa, ob, oc, od, stream);
}

int pf_step2_f32(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm, const float* a,
                 const float* b, const float* c, const float* d, float* oa, float* ob,
                 float* oc, float* od, void* stream) {
  return stepk_t<float>(2, ctx, nx, ny, bc, prm, a, b, c, d, oa, ob, oc, od, stream);
}

int pf_step3_f64(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm, const double* a,
                 const double* b, const double* c, const double* d, double* oa, double* ob,
                 double* oc, double* od, void* stream) {
  return stepk_t<double>(3, ctx, nx, ny, bc, prm, a, b, c, d, oa, ob, oc, od, stream);
}

int pf_step3_f32(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm, const float* a,
                 const float* b, const float* c, const float* d, float* oa, float* ob,
                 float* oc, float* od, void* stream) {
  return stepk_t<float>(3, ctx, nx, ny, bc, prm, a, b, c, d, oa, ob, oc, od, stream);
}

int pf_step4_f64(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm, const double* a,
                 const double* b, const double* c, const double* d, double* oa, double* ob,
                 double* oc, double* od, void* stream) {
  return stepk_t<double>(4, ctx, nx, ny, bc, prm, a, b, c, d, oa, ob, oc, od, stream);
}

int pf_step4_f32(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm, const float* a,
                 const float* b, const float* c, const float* d, float* oa, float* ob,
                 float* oc, float* od, void* stream) {
  return stepk_t<float>(4, ctx, nx, ny, bc, prm, a, b, c, d, oa, ob, oc, od, stream);
}

int pf_step_loop_f64(pf_ctx* ctx, int n, const pf_step_item* items, void* stream) {
  if (n < 0 || (n > 0 && !items)) return fail(PF_ERR_ARG, "bad item list");
  for (int i = 0; i < n; i++) {
    const pf_step_item& it = items[i];
    if (int r = step_t<double>(ctx, it.nx, it.ny, it.bc, it.prm, it.old_fields[0],
                               it.old_fields[1], it.old_fields[2], it.old_fields[3],
                               it.new_fields[0], it.new_fields[1], it.new_fields[2],
                               it.new_fields[3], nullptr, stream))
      return r;
  }
  return PF_OK;
}

static int stepk_loop(int k, pf_ctx* ctx, int n, const pf_step_item* items, void* stream) {
  if (n < 0 || (n > 0 && !items)) return fail(PF_ERR_ARG, "bad item list");
  for (int i = 0; i < n; i++) {
    const pf_step_item& it = items[i];
    if (int r = stepk_t<double>(k, ctx, it.nx, it.ny, it.bc, it.prm, it.old_fields[0],
                                it.old_fields[1], it.old_fields[2], it.old_fields[3],
                                it.new_fields[0], it.new_fields[1], it.new_fields[2],
                                it.new_fields[3], stream))
      return r;
  }
  return PF_OK;
}

int pf_step2_loop_f64(pf_ctx* ctx, int n, const pf_step_item* items, void* stream) {
  return stepk_loop(2, ctx, n, items, stream);
}

int pf_step3_loop_f64(pf_ctx* ctx, int n, const pf_step_item* items, void* stream) {
  return stepk_loop(3, ctx, n, items, stream);
}

int pf_step4_loop_f64(pf_ctx* ctx, int n, const pf_step_item* items, void* stream) {
  return stepk_loop(4, ctx, n, items, stream);
}

int pf_nucleation_f64(pf_ctx* ctx, int nx, int ny, double* a, double* b, double* c, double* d,
                      const pf_nuc_params* nuc, void* stream) {
  return nucleation_t<double>(ctx, nx, ny, a, b, c, d, nuc, stream);
}

int pf_nucleation_f32(pf_ctx* ctx, int nx, int ny, float* a, float* b, float* c, float* d,
                      const pf_nuc_params* nuc, void* stream) {
  return nucleation_t<float>(ctx, nx, ny, a, b, c, d, nuc, stream);
}

int pf_nucleation_status(pf_ctx* ctx, uint64_t* wide_events, int* overflow) {
  if (!ctx) return fail(PF_ERR_ARG, "null context");
  uint32_t h[4] = {0, 0, 0, 0};
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "synchronize");
  e = hipMemcpy(h, ctx->scratch.counters, sizeof(h), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "read nucleation counters");
  if (wide_events) *wide_events = h[2];
  if (overflow) *overflow = (int)h[1];
  if (h[1]) return fail(PF_ERR_OVERFLOW, "nucleation event list overflowed");
  return PF_OK;
}

int pf_summarize_f64(pf_ctx* ctx, int nx, int ny, const double* a, const double* b,
                     const double* c, const double* d, double* out5, void* stream) {
  return summarize_t<double>(ctx, nx, ny, a, b, c, d, out5, stream);
}

int pf_summarize_f32(pf_ctx* ctx, int nx, int ny, const float* a, const float* b,
                     const float* c, const float* d, double* out5, void* stream) {
  return summarize_t<float>(ctx, nx, ny, a, b, c, d, out5, stream);
}

int pf_energy_f64(pf_ctx* ctx, int nx, int ny, int bc, double dx, double dy,
                  const double* kappa, const double* omega, double Omega, const double* a,
                  const double* b, const double* c, const double* d, double* chem, double* grad,
                  double* out3, void* stream) {
  return energy_t<double>(ctx, nx, ny, bc, dx, dy, kappa, omega, Omega, a, b, c, d, chem, grad,
                          out3, stream);
}

int pf_energy_f32(pf_ctx* ctx, int nx, int ny, int bc, double dx, double dy,
                  const double* kappa, const double* omega, double Omega, const float* a,
                  const float* b, const float* c, const float* d, double* chem, double* grad,
                  double* out3, void* stream) {
  return energy_t<float>(ctx, nx, ny, bc, dx, dy, kappa, omega, Omega, a, b, c, d, chem, grad,
                         out3, stream);
}

int pf_dataviz_f64(int nx, int ny, const double* a, const double* b, const double* c,
                   const double* d, double* cNi, double* phi, void* stream) {
  if (int r = check_grid(nx, ny)) return r;
  if (int r = check_ptrs(a, b, c, d)) return r;
  if (!cNi || !phi) return fail(PF_ERR_ARG, "null output pointer");
  hipError_t e = pf::launch_dataviz(nx, ny, a, b, c, d, cNi, phi, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "dataviz kernel");
  return PF_OK;
}

int pf_thermo_eval(int what, int n, const double* in, double* out, void* stream) {
  if (what < 0 || what > 3) return fail(PF_ERR_ARG, "unknown function family");
  if (n < 0 || (n > 0 && (!in || !out))) return fail(PF_ERR_ARG, "bad sizes or pointers");
  hipError_t e = pf::launch_thermo_eval(what, n, in, out, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "thermo eval kernel");
  return PF_OK;
}

int pf_philox_eval(int n, const uint32_t* in, uint32_t* out, void* stream) {
  if (n < 0 || (n > 0 && (!in || !out))) return fail(PF_ERR_ARG, "bad sizes or pointers");
  hipError_t e = pf::launch_philox_eval(n, in, out, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "philox kernel");
  return PF_OK;
}

double pf_host_p(double x) { return pf::p(x); }
double pf_host_pPrime(double x) { return pf::pPrime(x); }
double pf_host_g(int phase, double xCr, double xNb) {
  if (phase < 0 || phase > 2) return NAN;
  return pf::g(phase, xCr, xNb);
}
double pf_host_D(int i, int j, double xCr, double xNb) {
  if (i < 0 || i > 1 || j < 0 || j > 1) return NAN;
  return pf::D(i, j, xCr, xNb);
}
void pf_host_fict(double xCr, double xNb, double pDel, double pGam, double pLav, double* out6) {
  double mu0, mu1;
  pf::kks_mu(xCr, xNb, pDel, pGam, pLav, mu0, mu1);
  for (int a = 0; a < 3; a++) {
    out6[2 * a] = pf::fict(a, 0, mu0, mu1);
    out6[2 * a + 1] = pf::fict(a, 1, mu0, mu1);
  }
}
double pf_host_xe(int phase, int comp) {
  if (phase < 0 || phase > 2 || comp < 0 || comp > 1) return NAN;
  return pf::XE[phase][comp];
}
void pf_host_nucleation_probability(double xCr, double xNb, double dG, double DCrCr,
                                    double DNbNb, double sigma, double Vatom, double n_gam,
                                    double dV, double dt, double* Rstar, double* P) {
  pf::nucleation_probability(xCr, xNb, dG, DCrCr, DNbNb, sigma, Vatom, n_gam, dV, dt, *Rstar,
                             *P);
}
double pf_host_max_effective_diffusivity(const double* DCr, const double* DNb) {
  // inside pure phase a: dx/dt = D A_gamma H_a lap(x); take the spectral
  // radius of the 2x2 product for a = gamma, delta, Laves
  double best = 0.0;
  for (int a = 0; a < 3; a++) {
    const double* Ag = pf::A[0];
    const double* Ha = pf::H[a];
    // K = A_gamma H_a
    const double K00 = Ag[0] * Ha[0] + Ag[1] * Ha[1], K01 = Ag[0] * Ha[1] + Ag[1] * Ha[2];
    const double K10 = Ag[1] * Ha[0] + Ag[2] * Ha[1], K11 = Ag[1] * Ha[1] + Ag[2] * Ha[2];
    // E = D K
    const double E00 = DCr[0] * K00 + DCr[1] * K10, E01 = DCr[0] * K01 + DCr[1] * K11;
    const double E10 = DNb[0] * K00 + DNb[1] * K10, E11 = DNb[0] * K01 + DNb[1] * K11;
    const double tr = E00 + E11, det = E00 * E11 - E01 * E10;
    const double disc = tr * tr / 4.0 - det;
    double lam;
    if (disc >= 0.0) {
      const double s = sqrt(disc);
      lam = fmax(fabs(tr / 2.0 + s), fabs(tr / 2.0 - s));
    } else {
      lam = sqrt(fabs(det));  // complex pair: modulus
    }
    if (lam > best) best = lam;
  }
  return best;
}
int pf_host_step_coefficients(const pf_params* prm, double* out22) {
  if (!prm || !out22) return fail(PF_ERR_ARG, "null parameters or output");
  const pf::StepParams P = pf::make_step(prm->dt, prm->DCr, prm->DNb, prm->kappa, prm->omega,
                                         prm->Lmob, prm->alpha, prm->mask);
  double* o = out22;
  for (int i = 0; i < 3; i++) *o++ = P.qCr[i];
  for (int i = 0; i < 3; i++) *o++ = P.qNb[i];
  for (int a = 0; a < 2; a++)
    for (int i = 0; i < 5; i++) *o++ = P.g[a][i];
  for (int a = 0; a < 2; a++) *o++ = P.w[a];
  for (int a = 0; a < 2; a++) *o++ = P.c[a];
  for (int a = 0; a < 2; a++) *o++ = P.k[a];
  return PF_OK;
}

}  // extern "C"
