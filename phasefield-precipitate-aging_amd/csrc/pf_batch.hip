// pf_batch.hip -- many independent simulations of uneven size advanced by one
// persistent launch per one, two or three steps.
//
// Layout: one device pool holds every job's two buffer sets (ping-pong), each
// set four planes (x_Cr, x_Nb, phi_delta, phi_Laves) of pitch x (Ny + 2) cells
// with the ghost ring, 256-byte aligned.  A device job table describes each job
// (pool offsets, size, tiling, kinetic parameters, step counter, target) and a
// wave table maps every strip task (job, 58-column strip, row chunk) of the
// current mix to its job.  The step kernel's grid is fixed (the GPU's resident
// capacity), so a captured hipGraph of several launches stays valid when jobs
// are admitted or retired: only the device tables change between replays.
//
// Per launch: step kernel (on every active job's strips, reading the job's
// current set and writing the other: pf::marchk when the job advances two or
// three steps this launch, pf::march for one), then the nucleation kernel
// (jobs whose schedule is due on the last step of the launch; centre-only
// embeds), then a one-block kernel that advances every active job's step
// counter, retires jobs that reached their target and decides each job's next
// advance.  A job advances k steps in a launch (k up to the batch's steps per
// launch) when k steps remain and no nucleation step falls before the last of
// them.
// A job's arithmetic and tiling depend only on the job, so its result is the
// same bits alone or in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "pf_march.h"
#include "pfpa.h"

namespace pf {

struct BatchJob {
  uint64_t off[2];       // element offsets of buffer sets 0 and 1 in the pool
  uint64_t plane;        // elements per field plane
  uint64_t step;         // steps done
  uint64_t target;       // steps to run
  int Nx, Ny, pitch, bc;
  int strips, chunks, rows, wave0;
  int active;            // 1 while step < target
  int nuc_every;         // 0 = no nucleation
  int parity;            // buffer set holding the current fields
  int adv;               // steps this launch advances the job (1, 2 or 3)
  int max_adv;           // steps per launch of the batch (1 where the grid is too small)
  int held;              // 1: the job takes no waves and does not advance (pf_batch_hold)
  StepParams P;
  NucParams N;
};

struct BatchHeader {
  int total_waves;
  int njobs;
  unsigned int pad;
  unsigned int nuc_counters[4];  // [0] deferred events, [1] overflow flag, [2] wide events
  int any_due;                   // some active job has a nucleation step in this launch
  // per-XCD task counters of the step kernel (element [x][0]; zero between
  // launches), one 128-byte line each: on one shared line the eight XCDs'
  // atomics contend and the FP64 batch ran 1.4% slower
  unsigned int next_task[8][32];
};

// the last step of this launch is one of the job's nucleation steps
__host__ __device__ __forceinline__ bool nuc_due(const BatchJob& J) {
  return J.active && !J.held && J.nuc_every > 0 &&
         (J.step + (uint64_t)J.adv - 1) % (uint64_t)J.nuc_every == 0;
}

constexpr int BATCH_RB = 16;  // reduction blocks per job

// persistent distribution: XCD x owns the contiguous wave range [x*T/8, (x+1)*T/8)
// = [first, last); stride = the XCD's wave count, start = this wave's slot in it
__device__ __forceinline__ void xcd_range(int total, int& first, int& last, int& stride,
                                          int& start) {
  const int x = blockIdx.x & 7;
  const int per_xcd_blocks = gridDim.x >> 3;
  const int lb = blockIdx.x >> 3;
  first = (int)(((long)total * x) >> 3);
  last = (int)(((long)total * (x + 1)) >> 3);
  stride = per_xcd_blocks * WPB;
  start = first + lb * WPB + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
}

// XCD x (the hardware places block b on XCD b % 8) works through the
// contiguous task range [x*T/8, (x+1)*T/8) in order.  Each of the XCD's
// waves takes one of the range's first tasks directly; after that a wave
// claims its next task from the XCD's counter (one atomic, by lane 0), so
// waves that draw short tasks take more of them.  The claim for the next
// task is issued before the current task runs and read after it; waves with
// no first task exit without touching the counter.
template <typename T, int DEPTH>
__global__ __launch_bounds__(WAVE * WPB) void batch_step_kernel(const BatchJob* __restrict__ jobs,
                                                                const int* __restrict__ wave_job,
                                                                BatchHeader* __restrict__ hdr,
                                                                T* __restrict__ pool) {
  int first, last, nwaves, w;
  xcd_range(hdr->total_waves, first, last, nwaves, w);
  unsigned int* ctr = &hdr->next_task[blockIdx.x & 7][0];
  const bool lane0 = (threadIdx.x & (WAVE - 1)) == 0;
  const NucParams N{};
  const Scratch S{};
  while (w < last) {
    unsigned int next = 0;
    if (lane0) next = atomicAdd(ctr, 1u);
    const int j = __builtin_amdgcn_readfirstlane(wave_job[w]);
    const BatchJob& J = jobs[j];
    if (!J.active) {
      w = first + nwaves + (int)__builtin_amdgcn_readfirstlane(next);
      continue;
    }
    const int local = w - J.wave0;
    const int strip = local % J.strips, chunk = local / J.strips;
    const int par = J.parity;
    const T* in = pool + J.off[par];
    T* out = pool + J.off[par ^ 1];
    const uint64_t pl = J.plane;
#define PF_BATCH_MARCH(BCV)                                                                  \
  if (J.adv == 3)                                                                            \
    marchk<T, BCV, 3, DEPTH, STRIP3>(in, in + pl, in + 2 * pl, in + 3 * pl, out, out + pl,     \
                                     out + 2 * pl, out + 3 * pl, J.Nx, J.Ny, J.pitch, strip,   \
                                     chunk, J.rows, J.P);                                      \
  else if (J.adv == 2)                                                                       \
    marchk<T, BCV, 2, DEPTH, STRIP3>(in, in + pl, in + 2 * pl, in + 3 * pl, out, out + pl,     \
                                     out + 2 * pl, out + 3 * pl, J.Nx, J.Ny, J.pitch, strip,   \
                                     chunk, J.rows, J.P);                                      \
  else                                                                                       \
    march<T, BCV, false, false, DEPTH, STRIP3>(in, in + pl, in + 2 * pl, in + 3 * pl, out,     \
                                               out + pl, out + 2 * pl, out + 3 * pl, J.Nx,    \
                                               J.Ny, J.pitch, strip, chunk, J.rows, J.P, N, S)
    if (J.bc == 0) {
      PF_BATCH_MARCH(0);
    } else {
      PF_BATCH_MARCH(1);
    }
#undef PF_BATCH_MARCH
    w = first + nwaves + (int)__builtin_amdgcn_readfirstlane(next);
  }
}

// Nucleation on the new fields of the jobs whose schedule is due on the last
// step of this launch (src/main.cpp:307-316 schedule, centre-only embeds);
// then batch_advance_kernel advances the step counters and picks each job's
// next advance (a multi-step launch never has a nucleation step before its
// last step).
__host__ __device__ __forceinline__ int next_adv(const BatchJob& J) {
  int k = 1;
  while (k < J.max_adv && J.step + k + 1 <= J.target &&
         !(J.nuc_every > 0 && (J.step + k - 1) % (uint64_t)J.nuc_every == 0))
    k++;
  return k;
}

template <typename T>
__global__ __launch_bounds__(WAVE * WPB) void batch_nuc_kernel(
    BatchJob* __restrict__ jobs, const int* __restrict__ wave_job, BatchHeader* __restrict__ hdr,
    T* __restrict__ pool) {
  int first, last, stride, w;
  xcd_range(hdr->total_waves, first, last, stride, w);
  const int lane = threadIdx.x & (WAVE - 1);
  // launches without a due job (all but one in nuc_every / steps per launch)
  // skip the scan of the wave table
  if (!hdr->any_due) return;
  for (; w < last; w += stride) {
    const int j = __builtin_amdgcn_readfirstlane(wave_job[w]);
    const BatchJob& J = jobs[j];
    const uint64_t s_last = J.step + J.adv - 1;  // the step whose output this is
    if (!nuc_due(J)) continue;
    const int local = w - J.wave0;
    const int strip = local % J.strips, chunk = local / J.strips;
    const int gx = strip * STRIP3 + lane - 1;
    if (lane < 1 || lane > STRIP3 || gx >= J.Nx) continue;
    T* out = pool + J.off[J.parity ^ 1];
    const uint64_t pl = J.plane;
    NucParams N = J.N;
    N.call = s_last / (uint64_t)J.nuc_every;
    const int y0 = chunk * J.rows, y1 = min(y0 + J.rows, J.Ny);
    for (int y = y0; y < y1; y++) {
      const size_t o = (size_t)(1 + y) * J.pitch + 1 + gx;
      // absolute compositions (the job's arrays hold x - xoff)
      double xCr = (double)out[o] + J.P.xoff[0], xNb = (double)out[o + pl] + J.P.xoff[1];
      double pd = (double)out[o + 2 * pl], plv = (double)out[o + 3 * pl];
      const double c0 = xCr, c1 = xNb, c2 = pd, c3 = plv;
      Scratch S{};
      S.counters = hdr->nuc_counters;  // wide events are counted, not applied, in batch mode
      S.event_capacity = 0;
      nucleate_cell(N, (uint32_t)((1 + y) * J.pitch + 1 + gx), xCr, xNb, pd, plv, S);
      if (xCr != c0 || xNb != c1 || pd != c2 || plv != c3) {
        xCr -= J.P.xoff[0];
        xNb -= J.P.xoff[1];
        store4(out, out + pl, out + 2 * pl, out + 3 * pl, o, xCr, xNb, pd, plv);
        if (J.bc == 0)
          store_ghost_images<T, 0>(out, out + pl, out + 2 * pl, out + 3 * pl, gx, y, J.Nx, J.Ny,
                                   J.pitch, xCr, xNb, pd, plv);
        else
          store_ghost_images<T, 1>(out, out + pl, out + 2 * pl, out + 3 * pl, gx, y, J.Nx, J.Ny,
                                   J.pitch, xCr, xNb, pd, plv);
      }
    }
  }
}

// One block after the nucleation kernel (stream order): advance every active
// job by this launch's steps, flip its buffer parity, retire it at its target,
// decide its next advance, and record whether any job is due at the next
// launch.
__global__ __launch_bounds__(256) void batch_advance_kernel(BatchJob* __restrict__ jobs,
                                                            BatchHeader* __restrict__ hdr) {
  int due = 0;
  for (int j = threadIdx.x; j < hdr->njobs; j += blockDim.x) {
    BatchJob& J = jobs[j];
    if (J.active && !J.held) {
      J.step += J.adv;
      J.parity ^= 1;
      if (J.step >= J.target) J.active = 0;
      J.adv = next_adv(J);
      due |= nuc_due(J) ? 1 : 0;
    }
  }
  due = __syncthreads_or(due);
  if (threadIdx.x == 0) hdr->any_due = due;
  if (threadIdx.x < 8) hdr->next_task[threadIdx.x][0] = 0u;  // the next step launch claims from 0
}

// per-job deterministic compensated sums: block (j, r) reduces rows
// [r*Ny/RB, (r+1)*Ny/RB) of job j's current fields
struct KahanB {
  double s, c;
};
__device__ __forceinline__ void kb_add(KahanB& k, double x) {
  const double t = k.s + x;
  if (fabs(k.s) >= fabs(x))
    k.c += (k.s - t) + x;
  else
    k.c += (x - t) + k.s;
  k.s = t;
}
__device__ __forceinline__ void kb_merge(KahanB& a, const KahanB& b) {
  kb_add(a, b.s);
  a.c += b.c;
}

template <typename T>
__global__ __launch_bounds__(256) void batch_partial_kernel(const BatchJob* __restrict__ jobs,
                                                            const T* __restrict__ pool,
                                                            double* __restrict__ partials,
                                                            int job) {
  const BatchJob& J = jobs[job];
  const int r = blockIdx.x;
  const int ya = (int)(((long)J.Ny * r) / BATCH_RB), yb = (int)(((long)J.Ny * (r + 1)) / BATCH_RB);
  const T* f = pool + J.off[J.parity];
  const uint64_t pl = J.plane;
  KahanB acc[5] = {};
  const long n = (long)(yb - ya) * J.Nx;
  for (long i = threadIdx.x; i < n; i += 256) {
    const int y = ya + (int)(i / J.Nx), x = (int)(i % J.Nx);
    const size_t o = (size_t)(1 + y) * J.pitch + 1 + x;
    const double d = p((double)f[o + 2 * pl]), l = p((double)f[o + 3 * pl]);
    kb_add(acc[0], (double)f[o]);
    kb_add(acc[1], (double)f[o + pl]);
    kb_add(acc[2], 1.0 - d - l);
    kb_add(acc[3], d);
    kb_add(acc[4], l);
  }
  __shared__ KahanB sh[5][256];
  for (int k = 0; k < 5; k++) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 5; k++) kb_merge(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; k++) {
      partials[(size_t)r * 10 + 2 * k] = sh[k][0].s;
      partials[(size_t)r * 10 + 2 * k + 1] = sh[k][0].c;
    }
}

__global__ void batch_final_kernel(const BatchJob* __restrict__ jobs, const double* partials,
                                   int job, double* out5) {
  const int k = threadIdx.x;
  if (k >= 5) return;
  const BatchJob& J = jobs[job];
  KahanB a{0.0, 0.0};
  for (int r = 0; r < BATCH_RB; r++) kb_merge(a, KahanB{partials[r * 10 + 2 * k], partials[r * 10 + 2 * k + 1]});
  out5[k] = (a.s + a.c) / ((double)J.Nx * (double)J.Ny) + (k < 2 ? J.P.xoff[k] : 0.0);
}

}  // namespace pf

// ======================================================================= host

struct pf_batch {
  int storage;  // PF_STORAGE_F32 / PF_STORAGE_F64
  size_t elem;
  size_t pool_elems;
  void* pool = nullptr;
  int max_jobs = 0;
  int max_waves = 0;
  pf::BatchJob* d_jobs = nullptr;
  int* d_wave_job = nullptr;
  pf::BatchHeader* d_hdr = nullptr;
  double* d_partials = nullptr;
  double* d_out5 = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t upload = nullptr;  // admissions' field uploads (overlap the running launches)
  int grid_blocks = 0;
  int rows = 16, depth = 3;
  int launch_steps = 1;  // steps per launch (1, 2 or 3)
  // host mirror
  std::vector<pf::BatchJob> jobs;       // slot -> job (slots are reused)
  std::vector<int> slot_live;           // 1 if the slot holds a job
  std::vector<int> slot_on_device;      // 1 if the device table holds the slot's current job
  std::map<uint64_t, uint64_t> free_;   // offset -> size (elements)
  bool dirty = true;
  // graph
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0, graph_depth = 0;
};

namespace {

int bfail(int code, const std::string& m) {
  pf::set_error(m);
  return code;
}
int bhip(hipError_t e, const char* where) {
  return bfail(PF_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

size_t align_elems(size_t n, size_t elem) {
  const size_t a = 256 / elem;
  return (n + a - 1) / a * a;
}

size_t job_elems(int Nx, int Ny, size_t elem) {
  const size_t plane = align_elems((size_t)(Nx + 2) * (Ny + 2), elem);
  return 8 * plane;
}

bool pool_alloc(pf_batch* b, size_t n, uint64_t& off) {
  for (auto it = b->free_.begin(); it != b->free_.end(); ++it) {
    if (it->second >= n) {
      off = it->first;
      const uint64_t rest = it->second - n;
      const uint64_t roff = it->first + n;
      b->free_.erase(it);
      if (rest) b->free_[roff] = rest;
      return true;
    }
  }
  return false;
}

void pool_free(pf_batch* b, uint64_t off, uint64_t n) {
  auto it = b->free_.emplace(off, n).first;
  // merge with the next block
  auto nx = std::next(it);
  if (nx != b->free_.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    b->free_.erase(nx);
  }
  // merge with the previous block
  if (it != b->free_.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      b->free_.erase(it);
    }
  }
}

int sync_jobs_from_device(pf_batch* b) {
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return bhip(e, "batch synchronize");
  if (b->jobs.empty()) return PF_OK;
  std::vector<pf::BatchJob> dev(b->jobs.size());
  e = hipMemcpy(dev.data(), b->d_jobs, dev.size() * sizeof(pf::BatchJob), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return bhip(e, "read job table");
  for (size_t j = 0; j < dev.size(); j++)
    if (b->slot_live[j] && b->slot_on_device[j]) {
      b->jobs[j].step = dev[j].step;
      b->jobs[j].active = dev[j].active;
      b->jobs[j].parity = dev[j].parity;
      b->jobs[j].adv = dev[j].adv;
    }
  return PF_OK;
}

template <typename T>
hipError_t launch_batch_step(pf_batch* b, hipStream_t st) {
  T* pool = (T*)b->pool;
  const dim3 grid(b->grid_blocks), block(pf::WAVE * pf::WPB);
  switch (b->depth) {
    case 2:
      hipLaunchKernelGGL((pf::batch_step_kernel<T, 2>), grid, block, 0, st, b->d_jobs,
                         b->d_wave_job, b->d_hdr, pool);
      break;
    case 6:
      hipLaunchKernelGGL((pf::batch_step_kernel<T, 6>), grid, block, 0, st, b->d_jobs,
                         b->d_wave_job, b->d_hdr, pool);
      break;
    default:
      hipLaunchKernelGGL((pf::batch_step_kernel<T, 3>), grid, block, 0, st, b->d_jobs,
                         b->d_wave_job, b->d_hdr, pool);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(pf::batch_nuc_kernel<T>, grid, block, 0, st, b->d_jobs, b->d_wave_job,
                     b->d_hdr, pool);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(pf::batch_advance_kernel, dim3(1), dim3(256), 0, st, b->d_jobs, b->d_hdr);
  return hipGetLastError();
}

hipError_t enqueue_step(pf_batch* b, hipStream_t st) {
  return b->storage == PF_STORAGE_F32 ? launch_batch_step<float>(b, st)
                                      : launch_batch_step<double>(b, st);
}

}  // namespace

extern "C" {

size_t pf_batch_job_bytes(int storage, int Nx, int Ny) {
  const size_t elem = storage == PF_STORAGE_F32 ? 4 : 8;
  return job_elems(Nx, Ny, elem) * elem;
}

int pf_batch_create(pf_batch** out, int storage, size_t pool_bytes, int max_jobs) {
  if (!out) return bfail(PF_ERR_ARG, "null output");
  *out = nullptr;
  if (storage != PF_STORAGE_F32 && storage != PF_STORAGE_F64)
    return bfail(PF_ERR_ARG, "storage must be PF_STORAGE_F32 or PF_STORAGE_F64");
  if (max_jobs < 1 || max_jobs > (1 << 20)) return bfail(PF_ERR_ARG, "max_jobs out of range");
  if (pool_bytes < (1u << 20)) return bfail(PF_ERR_ARG, "pool must be at least 1 MiB");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return bfail(PF_ERR_NO_DEVICE, "no HIP device visible");
  pf_batch* b = new pf_batch();
  b->storage = storage;
  b->elem = storage == PF_STORAGE_F32 ? 4 : 8;
  b->pool_elems = pool_bytes / b->elem;
  b->max_jobs = max_jobs;
  // every wave covers >= 58 x 4 cells of one job; jobs add at most one
  // partial strip and chunk per dimension
  const size_t cells = b->pool_elems / 8;
  b->max_waves = (int)std::min<size_t>((size_t)1 << 28,
                                       cells / (pf::STRIP3 * 4) + (size_t)max_jobs * 64);
  hipError_t e;
  if ((e = hipMalloc(&b->pool, b->pool_elems * b->elem)) != hipSuccess ||
      (e = hipMalloc((void**)&b->d_jobs, (size_t)max_jobs * sizeof(pf::BatchJob))) != hipSuccess ||
      (e = hipMalloc((void**)&b->d_wave_job, (size_t)b->max_waves * sizeof(int))) != hipSuccess ||
      (e = hipMalloc((void**)&b->d_hdr, sizeof(pf::BatchHeader))) != hipSuccess ||
      (e = hipMalloc((void**)&b->d_partials, pf::BATCH_RB * 10 * sizeof(double))) != hipSuccess ||
      (e = hipMalloc((void**)&b->d_out5, 8 * sizeof(double))) != hipSuccess ||
      (e = hipStreamCreate(&b->stream)) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&b->upload, hipStreamNonBlocking)) != hipSuccess) {
    pf_batch_destroy(b);
    return bhip(e, "batch allocation");
  }
  pf::BatchHeader h{};
  if ((e = hipMemcpy(b->d_hdr, &h, sizeof(h), hipMemcpyHostToDevice)) != hipSuccess) {
    pf_batch_destroy(b);
    return bhip(e, "batch header");
  }
  b->free_[0] = b->pool_elems;
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (cus < 8) cus = 8;
  b->grid_blocks = (cus * 4 + 7) / 8 * 8;  // 4 blocks of 4 waves per CU, multiple of 8 XCDs
  *out = b;
  return PF_OK;
}

void pf_batch_destroy(pf_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->pool) (void)hipFree(b->pool);
  if (b->d_jobs) (void)hipFree(b->d_jobs);
  if (b->d_wave_job) (void)hipFree(b->d_wave_job);
  if (b->d_hdr) (void)hipFree(b->d_hdr);
  if (b->d_partials) (void)hipFree(b->d_partials);
  if (b->d_out5) (void)hipFree(b->d_out5);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->upload) (void)hipStreamDestroy(b->upload);
  delete b;
}

size_t pf_batch_free_bytes(pf_batch* b) {
  if (!b) return 0;
  size_t best = 0;
  for (auto& kv : b->free_) best = std::max<size_t>(best, kv.second);
  return best * b->elem;
}

int pf_batch_set_tile(pf_batch* b, int rows, int depth) {
  if (!b) return bfail(PF_ERR_ARG, "null batch");
  if (rows < 1 || rows > 1 << 16) return bfail(PF_ERR_ARG, "rows out of range");
  if (depth != 2 && depth != 3 && depth != 6) return bfail(PF_ERR_ARG, "depth must be 2, 3 or 6");
  if (int r = sync_jobs_from_device(b)) return r;
  b->rows = rows;
  b->depth = depth;
  b->dirty = true;
  return PF_OK;
}

int pf_batch_set_launch_steps(pf_batch* b, int steps) {
  if (!b) return bfail(PF_ERR_ARG, "null batch");
  if (steps < 1 || steps > 3) return bfail(PF_ERR_ARG, "steps per launch must be 1, 2 or 3");
  if (int r = sync_jobs_from_device(b)) return r;
  b->launch_steps = steps;
  b->dirty = true;
  return PF_OK;
}

int pf_batch_add(pf_batch* b, const pf_job_desc* d, const void* xCr, const void* xNb,
                 const void* pd, const void* pl, int* job_id) {
  if (!b || !d || !job_id) return bfail(PF_ERR_ARG, "null argument");
  if (!xCr || !xNb || !pd || !pl) return bfail(PF_ERR_ARG, "null field pointer");
  if (d->Nx < 1 || d->Ny < 1 || (long long)(d->Nx + 2) * (d->Ny + 2) >= (1LL << 31))
    return bfail(PF_ERR_ARG, "bad grid size");
  if (d->bc != PF_BC_NOFLUX && d->bc != PF_BC_PERIODIC) return bfail(PF_ERR_ARG, "bad boundary mode");
  for (int c = 0; c < 2; c++) {
    if (!(d->x_offset[c] - d->x_offset[c] == 0.0))
      return bfail(PF_ERR_ARG, "x_offset must be finite");
    if (b->storage == PF_STORAGE_F64 && d->x_offset[c] != 0.0)
      return bfail(PF_ERR_ARG, "x_offset applies to FP32 storage only");
  }
  const double* m = d->prm.mask;
  if (!pf::symmetric_five_point(m))
    return bfail(PF_ERR_ARG,
                 "batched form supports the symmetric five-point Laplacian only (zero corners, "
                 "m[0][1] = m[2][1] != 0, m[1][0] = m[1][2])");
  if (d->nucleation) {
    const double R_scan = 1.75e-9 / (d->dx * (d->Nx + 2));
    if (!(R_scan < 1.0))
      return bfail(PF_ERR_ARG, "nucleation scan radius >= 1 cell: domain too small for batch mode");
    if (d->nuc_every < 1) return bfail(PF_ERR_ARG, "nuc_every must be >= 1");
  }
  // No wait for the launches in flight: the new job's slot and pool region
  // are unused by them (a slot or region is freed only by pf_batch_remove,
  // which waits for the stream, and the tables the launches read change only
  // at the next pf_batch_commit, which waits for the stream too), so its
  // fields upload while they run.
  int slot = -1;
  for (size_t j = 0; j < b->jobs.size(); j++)
    if (!b->slot_live[j]) {
      slot = (int)j;
      break;
    }
  if (slot < 0) {
    if ((int)b->jobs.size() >= b->max_jobs) return bfail(PF_ERR_ARG, "job table full");
    slot = (int)b->jobs.size();
    b->jobs.emplace_back();
    b->slot_live.push_back(0);
    b->slot_on_device.push_back(0);
  }
  const size_t n = job_elems(d->Nx, d->Ny, b->elem);
  uint64_t off;
  if (!pool_alloc(b, n, off)) return bfail(PF_ERR_OVERFLOW, "not enough pool memory for the job");
  pf::BatchJob J{};
  J.plane = n / 8;
  J.off[0] = off;
  J.off[1] = off + 4 * J.plane;
  J.step = 0;
  J.parity = 0;
  J.target = d->steps;
  J.Nx = d->Nx;
  J.Ny = d->Ny;
  J.pitch = d->Nx + 2;
  J.bc = d->bc;
  J.active = d->steps > 0 ? 1 : 0;
  J.nuc_every = d->nucleation ? d->nuc_every : 0;
  J.P = pf::make_step(d->prm.dt, d->prm.DCr, d->prm.DNb, d->prm.kappa, d->prm.omega,
                      d->prm.Lmob, d->prm.alpha, d->prm.mask, d->x_offset);
  J.N = pf::make_nuc(d->sigma_del, d->sigma_lav, d->unit_a, d->dx, d->dx, d->dx,
                     (double)(d->nuc_every > 0 ? d->nuc_every : 1) * d->prm.dt, d->seed, 0,
                     d->instance, d->Nx + 2, d->Ny + 2);
  // upload the initial fields into set 0 (ghost ring included)
  const size_t bytes = (size_t)(d->Nx + 2) * (d->Ny + 2) * b->elem;
  const void* src[4] = {xCr, xNb, pd, pl};
  char* base = (char*)b->pool;
  hipError_t e;
  for (int k = 0; k < 4; k++) {
    e = hipMemcpyAsync(base + (J.off[0] + k * J.plane) * b->elem, src[k], bytes,
                       hipMemcpyHostToDevice, b->upload);
    if (e != hipSuccess) {
      pool_free(b, off, n);
      return bhip(e, "upload job fields");
    }
  }
  e = hipStreamSynchronize(b->upload);
  if (e != hipSuccess) return bhip(e, "upload job fields");
  b->jobs[slot] = J;
  b->slot_live[slot] = 1;
  b->slot_on_device[slot] = 0;
  b->dirty = true;
  *job_id = slot;
  return PF_OK;
}

int pf_batch_remove(pf_batch* b, int job_id) {
  if (!b) return bfail(PF_ERR_ARG, "null batch");
  if (job_id < 0 || job_id >= (int)b->jobs.size() || !b->slot_live[job_id])
    return bfail(PF_ERR_ARG, "no such job");
  if (int r = sync_jobs_from_device(b)) return r;
  pf::BatchJob& J = b->jobs[job_id];
  pool_free(b, J.off[0], 8 * J.plane);
  J.active = 0;
  b->slot_live[job_id] = 0;
  b->slot_on_device[job_id] = 0;
  b->dirty = true;
  return PF_OK;
}

int pf_batch_commit(pf_batch* b) {
  if (!b) return bfail(PF_ERR_ARG, "null batch");
  if (!b->dirty) return PF_OK;
  if (int r = sync_jobs_from_device(b)) return r;
  std::vector<int> wave_job;
  for (size_t j = 0; j < b->jobs.size(); j++) {
    pf::BatchJob& J = b->jobs[j];
    if (!b->slot_live[j]) {
      J.active = 0;
      J.wave0 = 0;
      J.strips = J.chunks = 0;
      continue;
    }
    J.rows = b->rows;
    J.strips = (J.Nx + pf::STRIP3 - 1) / pf::STRIP3;
    J.chunks = (J.Ny + J.rows - 1) / J.rows;
    J.max_adv = (J.Nx >= 3 && J.Ny >= 3) ? b->launch_steps : 1;
    J.adv = J.active ? pf::next_adv(J) : 1;
    J.wave0 = (int)wave_job.size();
    if (!J.active || J.held) {
      J.wave0 = 0;
      continue;  // finished and held jobs take no waves
    }
    const long nw = (long)J.strips * J.chunks;
    if ((long)wave_job.size() + nw > b->max_waves) return bfail(PF_ERR_OVERFLOW, "wave table full");
    wave_job.insert(wave_job.end(), (size_t)nw, (int)j);
  }
  hipError_t e;
  if (!b->jobs.empty() &&
      (e = hipMemcpy(b->d_jobs, b->jobs.data(), b->jobs.size() * sizeof(pf::BatchJob),
                     hipMemcpyHostToDevice)) != hipSuccess)
    return bhip(e, "upload job table");
  if (!wave_job.empty() &&
      (e = hipMemcpy(b->d_wave_job, wave_job.data(), wave_job.size() * sizeof(int),
                     hipMemcpyHostToDevice)) != hipSuccess)
    return bhip(e, "upload wave table");
  pf::BatchHeader h;
  if ((e = hipMemcpy(&h, b->d_hdr, sizeof(h), hipMemcpyDeviceToHost)) != hipSuccess)
    return bhip(e, "read header");
  h.total_waves = (int)wave_job.size();
  h.njobs = (int)b->jobs.size();
  for (int x = 0; x < 8; x++) h.next_task[x][0] = 0u;
  h.pad = 0u;
  h.any_due = 0;
  for (size_t j = 0; j < b->jobs.size(); j++)
    if (b->slot_live[j] && pf::nuc_due(b->jobs[j])) h.any_due = 1;
  if ((e = hipMemcpy(b->d_hdr, &h, sizeof(h), hipMemcpyHostToDevice)) != hipSuccess)
    return bhip(e, "upload header");
  for (size_t j = 0; j < b->jobs.size(); j++) b->slot_on_device[j] = b->slot_live[j];
  b->dirty = false;
  return PF_OK;
}

int pf_batch_run(pf_batch* b, int launches, int per_graph) {
  if (!b) return bfail(PF_ERR_ARG, "null batch");
  if (launches < 0 || per_graph < 0) return bfail(PF_ERR_ARG, "negative launch count");
  if (int r = pf_batch_commit(b)) return r;
  hipError_t e;
  int done = 0;
  if (per_graph > 1) {
    if (!b->graph || b->graph_steps != per_graph || b->graph_depth != b->depth) {
      if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
      }
      hipGraph_t g;
      if ((e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return bhip(e, "begin capture");
      hipError_t le = hipSuccess;
      for (int s = 0; s < per_graph && le == hipSuccess; s++) le = enqueue_step(b, b->stream);
      e = hipStreamEndCapture(b->stream, &g);
      if (le != hipSuccess) return bhip(le, "capture step");
      if (e != hipSuccess) return bhip(e, "end capture");
      e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return bhip(e, "instantiate graph");
      b->graph_steps = per_graph;
      b->graph_depth = b->depth;
    }
    for (; done + per_graph <= launches; done += per_graph)
      if ((e = hipGraphLaunch(b->graph, b->stream)) != hipSuccess) return bhip(e, "graph launch");
  }
  for (; done < launches; done++)
    if ((e = enqueue_step(b, b->stream)) != hipSuccess) return bhip(e, "step launch");
  return PF_OK;
}

int pf_batch_sync(pf_batch* b) {
  if (!b) return bfail(PF_ERR_ARG, "null batch");
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return bhip(e, "batch synchronize");
  return PF_OK;
}

int pf_batch_poll(pf_batch* b, int job_id, uint64_t* steps_done, int* active) {
  if (!b) return bfail(PF_ERR_ARG, "null batch");
  if (job_id < 0 || job_id >= (int)b->jobs.size() || !b->slot_live[job_id])
    return bfail(PF_ERR_ARG, "no such job");
  if (int r = sync_jobs_from_device(b)) return r;
  if (steps_done) *steps_done = b->jobs[job_id].step;
  if (active) *active = b->jobs[job_id].active;
  return PF_OK;
}

int pf_batch_poll_all(pf_batch* b, int capacity, int* job_ids, uint64_t* steps_done, int* active,
                      int* count) {
  if (!b || !count) return bfail(PF_ERR_ARG, "null argument");
  if (int r = sync_jobs_from_device(b)) return r;
  int n = 0;
  for (size_t j = 0; j < b->jobs.size(); j++) {
    if (!b->slot_live[j]) continue;
    if (n < capacity) {
      if (job_ids) job_ids[n] = (int)j;
      if (steps_done) steps_done[n] = b->jobs[j].step;
      if (active) active[n] = b->jobs[j].active;
    }
    n++;
  }
  *count = n;
  return PF_OK;
}

int pf_batch_read(pf_batch* b, int job_id, void* xCr, void* xNb, void* pd, void* pl) {
  if (!b) return bfail(PF_ERR_ARG, "null batch");
  if (job_id < 0 || job_id >= (int)b->jobs.size() || !b->slot_live[job_id])
    return bfail(PF_ERR_ARG, "no such job");
  if (int r = sync_jobs_from_device(b)) return r;
  const pf::BatchJob& J = b->jobs[job_id];
  const size_t bytes = (size_t)(J.Nx + 2) * (J.Ny + 2) * b->elem;
  void* dst[4] = {xCr, xNb, pd, pl};
  const char* base = (const char*)b->pool + J.off[J.parity] * b->elem;
  for (int k = 0; k < 4; k++) {
    if (!dst[k]) continue;
    hipError_t e = hipMemcpy(dst[k], base + k * J.plane * b->elem, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return bhip(e, "read job fields");
  }
  return PF_OK;
}

int pf_batch_summarize(pf_batch* b, int job_id, double* out5) {
  if (!b || !out5) return bfail(PF_ERR_ARG, "null argument");
  if (job_id < 0 || job_id >= (int)b->jobs.size() || !b->slot_live[job_id])
    return bfail(PF_ERR_ARG, "no such job");
  if (int r = pf_batch_commit(b)) return r;
  if (int r = sync_jobs_from_device(b)) return r;
  if (b->storage == PF_STORAGE_F32)
    hipLaunchKernelGGL(pf::batch_partial_kernel<float>, dim3(pf::BATCH_RB), dim3(256), 0,
                       b->stream, b->d_jobs, (const float*)b->pool, b->d_partials, job_id);
  else
    hipLaunchKernelGGL(pf::batch_partial_kernel<double>, dim3(pf::BATCH_RB), dim3(256), 0,
                       b->stream, b->d_jobs, (const double*)b->pool, b->d_partials, job_id);
  hipLaunchKernelGGL(pf::batch_final_kernel, dim3(1), dim3(64), 0, b->stream, b->d_jobs,
                     b->d_partials, job_id, b->d_out5);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bhip(e, "batch summarize");
  if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return bhip(e, "batch summarize");
  if ((e = hipMemcpy(out5, b->d_out5, 5 * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
    return bhip(e, "read summary");
  return PF_OK;
}

int pf_batch_status(pf_batch* b, uint64_t* wide_events) {
  if (!b) return bfail(PF_ERR_ARG, "null batch");
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return bhip(e, "batch synchronize");
  pf::BatchHeader h;
  if ((e = hipMemcpy(&h, b->d_hdr, sizeof(h), hipMemcpyDeviceToHost)) != hipSuccess)
    return bhip(e, "read header");
  if (wide_events) *wide_events = h.nuc_counters[2];
  if (h.nuc_counters[2])
    return bfail(PF_ERR_ARG, "nucleation events wider than one cell occurred (not applied in "
                             "batch mode)");
  return PF_OK;
}

int pf_batch_hold(pf_batch* b, int job_id, int hold) {
  if (!b) return bfail(PF_ERR_ARG, "null batch");
  if (job_id < 0 || job_id >= (int)b->jobs.size() || !b->slot_live[job_id])
    return bfail(PF_ERR_ARG, "no such job");
  const int h = hold ? 1 : 0;
  if (b->jobs[job_id].held == h) return PF_OK;
  if (int r = sync_jobs_from_device(b)) return r;
  b->jobs[job_id].held = h;
  b->dirty = true;
  return PF_OK;
}

void* pf_batch_stream(pf_batch* b) { return b ? (void*)b->stream : nullptr; }

int pf_batch_waves(pf_batch* b) {
  if (!b) return 0;
  int total = 0;
  for (size_t j = 0; j < b->jobs.size(); j++)
    if (b->slot_live[j] && b->jobs[j].active && !b->jobs[j].held)
      total += ((b->jobs[j].Nx + pf::STRIP3 - 1) / pf::STRIP3) *
               ((b->jobs[j].Ny + b->rows - 1) / b->rows);
  return total;
}

}  // extern "C"
